"""Python mirror of MetalFFT's convolution / correlation classes over the
libmifft C-ABI.

    Convolver(kernel, fft_size)            Convolve.swift:8-87
    FusedConvolver(kernel, fft_size=4096,  FusedConvolver.swift:17-134
                   precision=Precision.float32)
    Correlator(fft_size)                   Correlate.swift:7-62
    cross_spectral(channels, fft_size,     CrossSpectral.swift:55-173
                   hop_size, window="hann")
    CrossSpectral(fft_size, channels)      (device form of the same)

``apply`` / ``cross`` / ``auto`` take host arrays and block, as the Swift
methods do.  ``apply_device`` / ``backward_device`` / ``cross_device`` run
batches of signals already on the GPU in one launch.
"""
from __future__ import annotations

import ctypes
import enum

import numpy as np

from . import _lib
from ._lib import FFTError
from .api import _dev_ptr, _stream_handle


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_lib.fptr)


class Precision(enum.IntEnum):
    """FusedConvolver.Precision (FusedConvolver.swift:27-36)."""
    float32 = 0
    float16Pure = 1
    float16Storage = 2
    float16Mixed = 3


class Convolver:
    """Overlap-style FFT FIR convolution; output length S + K - 1."""

    _fused = 0

    def __init__(self, kernel, fft_size: int, precision: Precision = Precision.float32):
        taps = _f32(kernel)
        self._cv = ctypes.c_void_p()
        _lib.check(_lib.load().mifft_convolver_create(_fp(taps), taps.size, int(fft_size),
                                                      self._fused, int(precision),
                                                      ctypes.byref(self._cv)))
        L = _lib.load()
        self.blockSize = int(L.mifft_convolver_block_size(self._cv))
        self.hopSize = int(L.mifft_convolver_hop_size(self._cv))
        self.kernelLen = int(L.mifft_convolver_kernel_len(self._cv))
        self.precision = Precision(precision)

    @staticmethod
    def recommended_fft_size(kernel_count: int):
        n = int(_lib.load().mifft_convolver_recommended_fft_size(int(kernel_count)))
        return None if n < 0 else n

    def apply(self, signal) -> np.ndarray:
        s = _f32(signal)
        out = np.empty(s.size + self.kernelLen - 1, dtype=np.float32)
        _lib.check(_lib.load().mifft_convolver_apply_host(self._cv, _fp(s), s.size, _fp(out)))
        return out

    def apply_device(self, signals, out, samples: int, count: int, stream=None) -> None:
        _lib.check(_lib.load().mifft_convolver_apply_device(
            self._cv, _dev_ptr(signals), int(samples), int(count), _dev_ptr(out),
            _stream_handle(stream)))

    def set_taps_device(self, taps, stream=None) -> None:
        _lib.check(_lib.load().mifft_convolver_set_taps_device(self._cv, _dev_ptr(taps),
                                                               _stream_handle(stream)))

    def backward_device(self, signals, grad_out, grad_signal, grad_taps, samples: int,
                        count: int, stream=None) -> None:
        _lib.check(_lib.load().mifft_convolver_backward_device(
            self._cv, _dev_ptr(signals), int(samples), int(count), _dev_ptr(grad_out),
            _dev_ptr(grad_signal), _dev_ptr(grad_taps), _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_cv", None) and self._cv.value:
            _lib.load().mifft_convolver_destroy(self._cv)
            self._cv = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FusedConvolver(Convolver):
    """Single-launch FFT -> multiply -> IFFT convolution at fftSize 4096."""

    supportedSize = 4096
    _fused = 1

    def __init__(self, kernel, fft_size: int = 4096, precision: Precision = Precision.float32):
        super().__init__(kernel, fft_size, precision)


class Correlator:
    """Circular cross/auto-correlation r[m] = sum_n a[(n+m) mod N] b[n]."""

    def __init__(self, fft_size: int):
        self._c = ctypes.c_void_p()
        _lib.check(_lib.load().mifft_correlator_create(int(fft_size), ctypes.byref(self._c)))
        self.fftSize = int(fft_size)

    def cross(self, a, b) -> np.ndarray:
        a, b = _f32(a), _f32(b)
        out = np.empty(self.fftSize, dtype=np.float32)
        _lib.check(_lib.load().mifft_correlator_cross_host(self._c, _fp(a), a.size, _fp(b), b.size,
                                                           _fp(out)))
        return out

    def auto(self, signal) -> np.ndarray:
        s = _f32(signal)
        out = np.empty(self.fftSize, dtype=np.float32)
        _lib.check(_lib.load().mifft_correlator_auto_host(self._c, _fp(s), s.size, _fp(out)))
        return out

    def cross_device(self, a, na: int, b, nb: int, count: int, out, stream=None) -> None:
        _lib.check(_lib.load().mifft_correlator_cross_device(
            self._c, _dev_ptr(a), int(na), _dev_ptr(b), int(nb), int(count), _dev_ptr(out),
            _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_c", None) and self._c.value:
            _lib.load().mifft_correlator_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def window_coefficients(kind: str, n: int) -> np.ndarray:
    """WindowType.coefficients (Sources/MetalFFT/DSP/Window.swift:14-23) for
    the vDSP windows it calls, from their published definitions (periodic,
    denominator N): hann 0.5 (1 - cos 2 pi n/N) (vDSP_HANN_NORM), hamming
    0.54 - 0.46 cos 2 pi n/N, blackman 0.42 - 0.5 cos 2 pi n/N + 0.08 cos
    4 pi n/N, rectangular 1."""
    k = np.arange(n, dtype=np.float64)
    if kind == "hann":
        w = 0.5 * (1.0 - np.cos(2.0 * np.pi * k / n))
    elif kind == "hamming":
        w = 0.54 - 0.46 * np.cos(2.0 * np.pi * k / n)
    elif kind == "blackman":
        w = 0.42 - 0.5 * np.cos(2.0 * np.pi * k / n) + 0.08 * np.cos(4.0 * np.pi * k / n)
    elif kind == "rectangular":
        w = np.ones(n)
    else:
        raise ValueError(f"unknown window {kind!r}")
    return w.astype(np.float32)


class CrossSpectralResult:
    """CrossSpectralResult (CrossSpectral.swift:11-39): power[C][N],
    crossSpectra[P][N] (complex), coherence[P][N]; welch_msc[P][N] is the
    coherence of the averaged spectra (no counterpart in the reference)."""

    def __init__(self, channels, fft_size, power, cross, coherence, welch_msc):
        self.channels = channels
        self.fftSize = fft_size
        self.power = power
        self.crossSpectra = cross
        self.coherence = coherence
        self.welch_msc = welch_msc

    @property
    def pairCount(self) -> int:
        return self.channels * (self.channels - 1) // 2

    def pairIndex(self, i: int, j: int) -> int:
        if not (0 <= i < j < self.channels):
            raise FFTError(10, "pairIndex requires i < j < channels")
        return sum(self.channels - 1 - a for a in range(i)) + (j - i - 1)


class CrossSpectral:
    """Device-side cross-spectral analyser over C channels of FFT size N
    (mifft_xspec_*): ``compute_device`` takes a (C, samples) float32 CUDA
    tensor and writes the averaged outputs into CUDA tensors."""

    def __init__(self, fft_size: int, channels: int):
        self._x = ctypes.c_void_p()
        _lib.check(_lib.load().mifft_xspec_create(int(fft_size), int(channels), ctypes.byref(self._x)))
        self.fftSize, self.channels = int(fft_size), int(channels)
        self.pairs = self.channels * (self.channels - 1) // 2

    def frames(self, samples: int, hop: int) -> int:
        return int(_lib.load().mifft_xspec_frames(self._x, int(samples), int(hop)))

    def compute_device(self, signals, samples: int, hop: int, window, power, cross, coherence,
                       welch_msc=None, stream=None) -> None:
        _lib.check(_lib.load().mifft_xspec_device(
            self._x, _dev_ptr(signals) if signals is not None else None, int(samples), int(hop),
            _dev_ptr(window) if window is not None else None, _dev_ptr(power),
            _dev_ptr(cross) if cross is not None else None,
            _dev_ptr(coherence) if coherence is not None else None,
            _dev_ptr(welch_msc) if welch_msc is not None else None, _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_x", None) and self._x.value:
            _lib.load().mifft_xspec_destroy(self._x)
            self._x = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cross_spectral(channels, fft_size: int, hop_size: int, sample_rate: float = 0.0,
                   window="hann") -> CrossSpectralResult:
    """PSD.crossSpectral(channels:fftSize:hopSize:sampleRate:window:)
    (CrossSpectral.swift:55-173) on host arrays (sample_rate labels bins only
    and is unused, as in the reference).  Raises FFTError
    (MIFFT_ERR_PRECONDITION) where the reference traps."""
    import torch
    sig = [np.asarray(c, dtype=np.float32).reshape(-1) for c in channels]
    if not sig or len(sig) > 16:
        raise FFTError(10, "crossSpectral requires 1-16 channels")
    if any(s.size != sig[0].size for s in sig):
        raise FFTError(10, "all channels must have the same length")
    C, S, N = len(sig), sig[0].size, int(fft_size)
    an = CrossSpectral(N, C)
    try:
        P = an.pairs
        x = torch.from_numpy(np.stack(sig)).cuda()
        w = None if window is None else torch.from_numpy(
            window_coefficients(window, N) if isinstance(window, str)
            else np.asarray(window, np.float32)).cuda()
        pw = torch.empty((C, N), dtype=torch.float32, device="cuda")
        cr = torch.empty((max(P, 1), N), dtype=torch.complex64, device="cuda")
        co = torch.empty((max(P, 1), N), dtype=torch.float32, device="cuda")
        ms = torch.empty((max(P, 1), N), dtype=torch.float32, device="cuda")
        an.compute_device(x, S, hop_size, w, pw, cr, co, ms)
        torch.cuda.synchronize()
        return CrossSpectralResult(C, N, pw.cpu().numpy(), cr[:P].cpu().numpy(),
                                   co[:P].cpu().numpy(), ms[:P].cpu().numpy())
    finally:
        an.close()


__all__ = ["Convolver", "FusedConvolver", "Correlator", "Precision", "FFTError", "CrossSpectral",
           "CrossSpectralResult", "cross_spectral", "window_coefficients"]
