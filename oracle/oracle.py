"""CPU oracle for the batched complex FFT path -- TEST INFRASTRUCTURE ONLY.

Two layers:

* ``cpu_*``: ctypes calls into ``libcpu_fft.so`` (oracle/cpu_fft.c), the FP32
  restatement of the reference's radix-4/8 Stockham, four-step, overlap-add
  convolution and correlation algorithms.  It is the "reference CPU path" the
  GPU FP32 results must match within 1e-6 relative L2, and the CPU baseline
  timed by bench.py.
* ``f64_*``: FP64 numpy ground truth (pocketfft in complex128, pinned against
  the brute-force DFT of cpu_fft.c and closed forms by tests/test_oracle.py).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  The product path (metalfft_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpu_fft.so")
_lib = None

SUPPORTED_SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)


def build(force: bool = False) -> str:
    """Compile oracle/cpu_fft.c with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_fft.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                 fp, fp, ctypes.c_int]
        L.oracle_fft.restype = ctypes.c_int
        L.oracle_convolve.argtypes = [fp, ctypes.c_int64, fp, ctypes.c_int, ctypes.c_int, fp]
        L.oracle_convolve.restype = ctypes.c_int
        L.oracle_correlate.argtypes = [fp, ctypes.c_int64, fp, ctypes.c_int64, ctypes.c_int, fp]
        L.oracle_correlate.restype = ctypes.c_int
        L.oracle_twiddle_table.argtypes = [ctypes.c_int, fp]
        L.oracle_twiddle_table.restype = None
        L.oracle_dft_f64.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.oracle_dft_f64.restype = None
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# ---------------------------------------------------------------- FP32 CPU path
def cpu_fft(x: np.ndarray, inverse: bool = False, radix8: bool = False,
            threads: int = 1) -> np.ndarray:
    """Batched FP32 FFT over the last axis (complex64), reference algorithm."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    out = np.empty_like(x)
    rc = lib().oracle_fft(n, batch, int(inverse), int(radix8), _fptr(x.view(np.float32)),
                          _fptr(out.view(np.float32)), threads)
    if rc != 0:
        raise ValueError(f"unsupported FFT size {n}")
    return out


def cpu_convolve(signal: np.ndarray, taps: np.ndarray, fft_size: int) -> np.ndarray:
    s = np.ascontiguousarray(signal, dtype=np.float32)
    t = np.ascontiguousarray(taps, dtype=np.float32)
    out = np.empty(s.size + t.size - 1, dtype=np.float32)
    rc = lib().oracle_convolve(_fptr(s), s.size, _fptr(t), t.size, fft_size, _fptr(out))
    if rc != 0:
        raise ValueError("invalid convolver configuration")
    return out


def cpu_correlate(a: np.ndarray, b: np.ndarray, fft_size: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.empty(fft_size, dtype=np.float32)
    rc = lib().oracle_correlate(_fptr(a), a.size, _fptr(b), b.size, fft_size, _fptr(out))
    if rc != 0:
        raise ValueError("unsupported FFT size")
    return out


def twiddle_table(n: int) -> np.ndarray:
    """W_N^k, k in [0, N): FP64 generation rounded once to FP32 (complex64)."""
    out = np.empty(n, dtype=np.complex64)
    lib().oracle_twiddle_table(n, _fptr(out.view(np.float32)))
    return out


def dft_brute_f64(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """O(N^2) FP64 DFT (no 1/N on inverse) from cpu_fft.c."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().oracle_dft_f64(x.size, x.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         int(inverse))
    return out


# ---------------------------------------------------------------- FP64 truth
def f64_fft(x: np.ndarray) -> np.ndarray:
    return np.fft.fft(np.asarray(x, dtype=np.complex128), axis=-1)


def f64_ifft(x: np.ndarray) -> np.ndarray:
    return np.fft.ifft(np.asarray(x, dtype=np.complex128), axis=-1)


def f64_convolve(signal: np.ndarray, taps: np.ndarray) -> np.ndarray:
    """Full linear convolution, length S + K - 1 (Convolve.swift:35-37)."""
    s = np.asarray(signal, dtype=np.float64)
    t = np.asarray(taps, dtype=np.float64)
    n = s.size + t.size - 1
    m = 1 << max(0, (n - 1).bit_length())
    return np.fft.irfft(np.fft.rfft(s, m) * np.fft.rfft(t, m), m)[:n]


def f64_correlate(a: np.ndarray, b: np.ndarray, fft_size: int) -> np.ndarray:
    """r[m] = sum_n a[(n+m) mod N] b[n] with truncation/zero-pad to N
    (Correlate.swift:33-61 as implemented)."""
    n = fft_size
    pa = np.zeros(n)
    pb = np.zeros(n)
    pa[: min(n, len(a))] = np.asarray(a, dtype=np.float64)[:n]
    pb[: min(n, len(b))] = np.asarray(b, dtype=np.float64)[:n]
    return np.real(np.fft.ifft(np.fft.fft(pa) * np.conj(np.fft.fft(pb))))


# ---------------------------------------------------------------- metrics
def rel_l2(a, b) -> float:
    """sqrt(sum|a-b|^2 / sum|b|^2), Tests/MetalFFTTests/MetalFFTTests.swift:190-198."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    num = float(np.sum(np.abs(a - b) ** 2))
    den = float(np.sum(np.abs(b) ** 2))
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a, np.complex128) - np.asarray(b, np.complex128))))


def f64_conv_grads(x: np.ndarray, taps: np.ndarray, gy: np.ndarray):
    """FP64 gradients of y = conv(x_s, h) (per row of x) given dL/dy:
    grad_x[s, n] = sum_k h[k] gy[s, n+k];  grad_h[k] = sum_s sum_n x[s, n] gy[s, n+k]."""
    x = np.asarray(x, np.float64)
    h = np.asarray(taps, np.float64)
    gy = np.asarray(gy, np.float64)
    S, K = x.shape[-1], h.size
    L = gy.shape[-1]
    m = 1 << (L + K).bit_length()
    G = np.fft.rfft(gy, m, axis=-1)
    # correlation c[n] = sum_k a[k] g[n + k]  <=>  irfft(G * conj(A))
    gx = np.fft.irfft(G * np.conj(np.fft.rfft(h, m)), m, axis=-1)[..., :S]
    gh = np.fft.irfft(G * np.conj(np.fft.rfft(x, m, axis=-1)), m, axis=-1)[..., :K]
    return gx, gh.reshape(-1, K).sum(axis=0)
