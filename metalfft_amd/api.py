"""Python mirror of MetalFFT's public FFT API over the libmifft C-ABI.

Names, argument meaning and error behaviour follow the Swift interface
(docs/MetalFFT.swiftinterface:217-247 in the reference):

    MetalFFT(size)                      MetalFFT.init(size:)           MetalFFT.swift:27-51
    .forward(x)                         forward(_:)                    MetalFFT.swift:69-85
    .forward_into(input, output)        forward(input:output:)         MetalFFT.swift:88-104
    .forward_batch(batch)               forward(batch:)                MetalFFT.swift:108-133
    .inverse(x)                         inverse(_:)                    MetalFFT.swift:56-65
    MultiChannelFFT(channels, size)     MultiChannelFFT                MultiChannelFFT.swift:19-54
    FFTError                            FFTError                       MetalContext.swift:6-30

Host arrays are numpy complex64 (interleaved re/im, the SIMD2<Float> layout).
``execute`` additionally runs on device buffers (torch tensors on the GPU),
asynchronously on a HIP stream, which the Swift API has no counterpart for.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import FORWARD, INVERSE, FFTError

SUPPORTED_SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)
# lengths beyond the reference's table, run as two-pass four-step transforms
FOURSTEP_SIZES = (32768, 65536, 131072, 262144, 524288, 1048576)


def _as_c64(x) -> np.ndarray:
    a = np.asarray(x)
    if a.dtype == np.complex64:
        return np.ascontiguousarray(a.reshape(-1))
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a.reshape(-1).astype(np.complex64))
    if a.ndim >= 1 and a.shape[-1] == 2 and a.dtype.kind == "f":
        # (N, 2) real pairs, like [SIMD2<Float>]
        return np.ascontiguousarray(a.astype(np.float32)).view(np.complex64).reshape(-1)
    return np.ascontiguousarray(a.reshape(-1).astype(np.complex64))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_lib.fptr)


def device_name() -> str:
    L = _lib.load()
    buf = ctypes.create_string_buffer(256)
    _lib.check(L.mifft_device_name(buf, 256))
    return buf.value.decode()


def kernel_variants(n: int) -> list[str]:
    L = _lib.load()
    cnt = L.mifft_kernel_variants(n, None, 0)
    arr = (ctypes.c_char_p * max(cnt, 1))()
    L.mifft_kernel_variants(n, arr, cnt)
    return [arr[i].decode() for i in range(cnt)]


def _stream_handle(stream) -> ctypes.c_void_p:
    if stream is None:
        try:
            import torch
            stream = torch.cuda.current_stream()
        except Exception:  # no torch: null stream
            return ctypes.c_void_p(0)
    h = getattr(stream, "cuda_stream", stream)
    return ctypes.c_void_p(int(h))


# Element types a device buffer may have, by what the entry point reads or
# writes there: real float32 samples, complex64 (or its float32 re/im view),
# complex FP16 (or its float16 re/im view).
F32 = ("torch.float32",)
C64 = ("torch.complex64", "torch.float32")
C32H = ("torch.complex32", "torch.float16")


def _dev_ptr(t, nbytes: int | None = None, dtypes=None, what: str = "buffer") -> int:
    """Device address of ``t`` for a *_device entry point.

    For a torch tensor: it must be on the GPU, contiguous, of one of
    ``dtypes`` and hold at least ``nbytes`` bytes, else FFTError
    (invalidArgument) is raised before anything is launched.  An integer is
    taken as a raw device address, unchecked."""
    if hasattr(t, "data_ptr"):
        if not t.is_cuda:
            raise FFTError(11, f"{what}: device entry point needs a GPU tensor")
        if not t.is_contiguous():
            raise FFTError(11, f"{what}: device tensor must be contiguous")
        if dtypes is not None and str(t.dtype) not in dtypes:
            raise FFTError(11, f"{what}: dtype {t.dtype} where {' or '.join(dtypes)} is needed")
        if nbytes is not None:
            have = t.numel() * t.element_size()
            if have < nbytes:
                raise FFTError(11, f"{what}: tensor holds {have} bytes, the call needs {nbytes}",
                               expected=int(nbytes), got=int(have))
        return int(t.data_ptr())
    return int(t)


def _count(v, what: str) -> int:
    v = int(v)
    if v < 0:
        raise FFTError(11, f"negative {what}")
    return v


class MetalFFT:
    """Batched complex FFT of one length: 64 ... 16384 (one workgroup per
    transform) or 32768 ... 2^20 (four-step, two launches), powers of two."""

    def __init__(self, size: int):
        L = _lib.load()
        self._plan = ctypes.c_void_p()
        _lib.check(L.mifft_plan_create(int(size), ctypes.byref(self._plan)))
        self.size = int(size)

    # -- plan control -------------------------------------------------------
    @property
    def kernel_name(self) -> str:
        return _lib.load().mifft_plan_kernel_name(self._plan).decode()

    def select_kernel(self, name: str) -> None:
        _lib.check(_lib.load().mifft_plan_select_kernel(self._plan, name.encode()))

    def close(self) -> None:
        if getattr(self, "_plan", None) and self._plan.value:
            _lib.load().mifft_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- host API (blocking, copies; reference semantics) -------------------
    def forward(self, x) -> np.ndarray:
        a = _as_c64(x)
        out = np.empty(self.size, dtype=np.complex64)
        _lib.check(_lib.load().mifft_forward_host(self._plan, _fp(a.view(np.float32)), a.size,
                                                  _fp(out.view(np.float32))))
        return out

    def forward_into(self, input, output: np.ndarray | None) -> np.ndarray:
        """forward(input:output:): writes into ``output`` when it already
        holds ``size`` complex64 values, else returns a new array (the
        reference resizes its inout array, MetalFFT.swift:95)."""
        a = _as_c64(input)
        if (output is None or not isinstance(output, np.ndarray) or output.dtype != np.complex64
                or output.size != self.size or not output.flags.c_contiguous):
            if a.size != self.size:  # same check order as the reference
                self.forward(a)
            output = np.empty(self.size, dtype=np.complex64)
        _lib.check(_lib.load().mifft_forward_host(self._plan, _fp(a.view(np.float32)), a.size,
                                                  _fp(output.view(np.float32))))
        return output

    def inverse(self, x) -> np.ndarray:
        a = _as_c64(x)
        out = np.empty(self.size, dtype=np.complex64)
        _lib.check(_lib.load().mifft_inverse_host(self._plan, _fp(a.view(np.float32)), a.size,
                                                  _fp(out.view(np.float32))))
        return out

    def _batch(self, batch: Sequence, fn) -> list[np.ndarray]:
        if len(batch) == 0:
            return []
        arrs = [_as_c64(b) for b in batch]
        outs = [np.empty(self.size, dtype=np.complex64) for _ in arrs]
        n = len(arrs)
        ins = (_lib.fptr * n)(*[_fp(a.view(np.float32)) for a in arrs])
        ous = (_lib.fptr * n)(*[_fp(o.view(np.float32)) for o in outs])
        cnt = (ctypes.c_int64 * n)(*[a.size for a in arrs])
        _lib.check(fn(self._plan, ins, cnt, n, ous))
        return outs

    def forward_batch(self, batch: Sequence) -> list[np.ndarray]:
        return self._batch(batch, _lib.load().mifft_forward_batch_host)

    def inverse_batch(self, batch: Sequence) -> list[np.ndarray]:
        return self._batch(batch, _lib.load().mifft_inverse_batch_host)

    # -- device API -----------------------------------------------------------
    def execute(self, src, dst, batch: int | None = None, direction: int = FORWARD,
                stream=None) -> None:
        """Transform ``batch`` contiguous length-``size`` complex64 signals
        already on the GPU (torch tensors or raw device addresses),
        asynchronously on ``stream`` (default: torch's current stream)."""
        if batch is None:
            batch = src.numel() // self.size
            if hasattr(src, "dtype") and not src.is_complex():
                batch = src.numel() // (2 * self.size)
        batch = _count(batch, "batch")
        need = 8 * self.size * batch
        _lib.check(_lib.load().mifft_execute_device(
            self._plan, _dev_ptr(src, need, C64, "src"), _dev_ptr(dst, need, C64, "dst"),
            batch, int(direction), _stream_handle(stream)))

    def forward_real(self, src, dst, batch: int, half: bool = False, stream=None) -> None:
        """Spectra of ``batch`` real float32 signals of length ``size`` on the
        GPU (mifft_forward_real_device): ``dst`` holds ``size`` complex64 bins
        per signal, or ``size // 2 + 1`` with ``half``."""
        batch = _count(batch, "batch")
        bins = self.size // 2 + 1 if half else self.size
        _lib.check(_lib.load().mifft_forward_real_device(
            self._plan, _dev_ptr(src, 4 * self.size * batch, F32, "src"),
            _dev_ptr(dst, 8 * bins * batch, C64, "dst"), batch, int(bool(half)),
            _stream_handle(stream)))

    def inverse_real(self, src, dst, batch: int, half: bool = False, stream=None) -> None:
        """Real signals from Hermitian spectra (mifft_inverse_real_device):
        ``src`` rows of ``size`` (or ``size // 2 + 1`` with ``half``)
        complex64, ``dst`` ``batch`` x ``size`` float32."""
        batch = _count(batch, "batch")
        bins = self.size // 2 + 1 if half else self.size
        _lib.check(_lib.load().mifft_inverse_real_device(
            self._plan, _dev_ptr(src, 8 * bins * batch, C64, "src"), int(bool(half)),
            _dev_ptr(dst, 4 * self.size * batch, F32, "dst"), batch, _stream_handle(stream)))

    def forwardReal(self, signal) -> np.ndarray:
        """Spectrum of one real host signal (all ``size`` bins), the result
        of forward() on the (x, 0) packing the reference's callers build."""
        import torch
        x = np.ascontiguousarray(np.asarray(signal, np.float32).reshape(-1))
        if x.size != self.size:
            raise FFTError(6, "Invalid input size", expected=self.size, got=x.size)
        dx = torch.from_numpy(x).cuda()
        y = torch.empty(self.size, dtype=torch.complex64, device="cuda")
        self.forward_real(dx, y, 1)
        torch.cuda.synchronize()
        return y.cpu().numpy()


class MultiChannelFFT:
    """Same-size FFTs over several channels in one launch per call
    (Sources/MetalFFT/FFT/MultiChannelFFT.swift:19-54); the inverse is one
    batched launch too instead of the reference's per-channel loop (:49-53)."""

    def __init__(self, channels: int, size: int):
        if channels <= 0:
            raise FFTError(10, "channels must be > 0")
        self.channels = int(channels)
        self.size = int(size)
        self._fft = MetalFFT(size)

    def _check(self, inputs):
        if len(inputs) != self.channels:
            raise FFTError(10, f"Expected {self.channels} channels, got {len(inputs)}")

    def forward(self, inputs: Sequence) -> list[np.ndarray]:
        self._check(inputs)
        return self._fft.forward_batch(inputs)

    def inverse(self, inputs: Sequence) -> list[np.ndarray]:
        self._check(inputs)
        return self._fft.inverse_batch(inputs)


RAGGED_DEFAULT = 0   # one launch, one workgroup per ticket (longest lengths first)
RAGGED_PER_SIZE = 1  # one launch per distinct length
RAGGED_GRAPH = 2     # hipGraph capture and replay (combines with the others)
RAGGED_QUEUE = 4     # one persistent launch fed by a device ticket counter
RAGGED_SPLIT = 8     # one launch per length >= 4096 + one persistent launch of short tiles


class RaggedFFT:
    """Transforms of different lengths (64 ... 16384) in one launch.

    ``lengths[i]`` complex values of transform i start at ``offsets[i]``
    (default: packed back to back).  ``execute`` runs all of them forward or
    inverse on device buffers; ``mode`` picks the launch structure (default:
    one launch of length-sorted tickets; RAGGED_PER_SIZE, RAGGED_QUEUE,
    RAGGED_SPLIT) and/or hipGraph replay (RAGGED_GRAPH)."""

    def __init__(self, lengths: Sequence[int], offsets: Sequence[int] | None = None):
        n = len(lengths)
        L = _lib.load()
        ln = (ctypes.c_int64 * max(n, 1))(*[int(v) for v in lengths])
        of = None
        if offsets is not None:
            of = (ctypes.c_int64 * max(n, 1))(*[int(v) for v in offsets])
        self._p = ctypes.c_void_p()
        _lib.check(L.mifft_ragged_plan_create(ln, of, n, ctypes.byref(self._p)))
        self.count = n
        self.extent = int(L.mifft_ragged_extent(self._p))
        self.elements = int(L.mifft_ragged_elements(self._p))

    def launches(self, mode: int = RAGGED_DEFAULT) -> int:
        return int(_lib.load().mifft_ragged_launches(self._p, int(mode)))

    def execute(self, src, dst, direction: int = FORWARD, mode: int = RAGGED_DEFAULT,
                stream=None) -> None:
        need = 8 * self.extent
        _lib.check(_lib.load().mifft_ragged_execute_device(
            self._p, _dev_ptr(src, need, C64, "src"), _dev_ptr(dst, need, C64, "dst"),
            int(direction), int(mode), _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_p", None) and self._p.value:
            _lib.load().mifft_ragged_plan_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PrecisionFFT:
    """N = 4096 batched FFT in an explicit arithmetic (mifft_precision) and
    device element type (0 complex64, 1 complex FP16) on device buffers."""

    FP32, FP16_MFMA, BF16_MFMA, FP16_MFMA_SPLIT, BF16_MFMA_SPLIT = range(5)
    C64, C32H = 0, 1

    def __init__(self, size: int, precision: int, io: int = 0):
        self._plan = ctypes.c_void_p()
        _lib.check(_lib.load().mifft_plan_create_ex(int(size), int(precision), int(io),
                                                    ctypes.byref(self._plan)))
        self.size, self.precision, self.io = int(size), int(precision), int(io)

    @property
    def kernel_name(self) -> str:
        return _lib.load().mifft_plan_kernel_name(self._plan).decode()

    def select_kernel(self, name: str) -> None:
        """Pins one of the plan's DFT-as-GEMM forms by name (mifft_plan_select_kernel)."""
        _lib.check(_lib.load().mifft_plan_select_kernel(self._plan, name.encode()))

    def execute(self, src, dst, batch: int, direction: int = FORWARD, stream=None) -> None:
        batch = _count(batch, "batch")
        kinds = C32H if self.io == self.C32H else C64
        need = (4 if self.io == self.C32H else 8) * self.size * batch
        _lib.check(_lib.load().mifft_execute_device(
            self._plan, _dev_ptr(src, need, kinds, "src"), _dev_ptr(dst, need, kinds, "dst"),
            batch, int(direction), _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_plan", None) and self._plan.value:
            _lib.load().mifft_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
