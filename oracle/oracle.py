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


# ---------------------------------------------------------------- SAR (FP64)
C_LIGHT = 299792458.0


def sar_default_params():
    """src/radar/sar_simulator.swift:22-67 (4096 x 4096)."""
    return dict(n_range=4096, n_azimuth=4096, bandwidth=100e6, pulse_duration=10e-6,
                sampling_rate=120e6, prf=1000.0, velocity=100.0, carrier_freq=10e9,
                center_range=20000.0)


def sar_simulate(p, targets, threads=8):
    """Raw echo (complex128, Na x Nr) of targets [(R0, x_t, complex amp)]."""
    L = lib()
    L.oracle_sar_simulate.argtypes = [ctypes.c_int64, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.oracle_sar_simulate.restype = None
    prm = np.array([p["bandwidth"], p["pulse_duration"], p["sampling_rate"], p["prf"],
                    p["velocity"], p["carrier_freq"], p["center_range"]], np.float64)
    t = np.array([[r, x, a.real, a.imag] for r, x, a in targets], np.float64).reshape(-1)
    out = np.zeros((p["n_azimuth"], p["n_range"]), np.complex128)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    L.oracle_sar_simulate(p["n_range"], p["n_azimuth"], dp(prm), len(targets), dp(t),
                          dp(out.view(np.float64)), threads)
    return out


def sar_chirp_reference(p):
    """src/radar/sar_simulator.swift:207-224 in FP64."""
    nr = p["n_range"]
    kr = p["bandwidth"] / p["pulse_duration"]
    ns = int(p["pulse_duration"] * p["sampling_rate"])
    half = ns // 2
    c = np.zeros(nr, np.complex128)
    i = np.arange(ns)
    t = (i - half) / p["sampling_rate"]
    idx = np.where(i < half, nr - half + i, i - half)
    ok = (idx >= 0) & (idx < nr)
    c[idx[ok]] = np.exp(1j * np.pi * kr * t[ok] ** 2)
    return c, ns


def sar_azimuth_filter(p, reference_formula=False):
    """H_a (Nr x Na) of src/radar/rda_kernels.metal:195-238 in FP64:
    exp(+i phi) (matched, conjugate of the echo's Doppler phase) or, with
    reference_formula, the kernel's literal exp(-i phi)."""
    nr, na = p["n_range"], p["n_azimuth"]
    lam = C_LIGHT / p["carrier_freq"]
    pixel = C_LIGHT / (2.0 * p["sampling_rate"])
    r0 = p["center_range"] + (np.arange(nr) - nr // 2) * pixel
    k = np.arange(na)
    fa = np.where(k < na // 2, k, k - na) / na * p["prf"]
    ratio = lam * fa / (2.0 * p["velocity"])
    d = np.sqrt(np.maximum(0.0, 1.0 - ratio ** 2))
    cyc = 2.0 * r0[:, None] * d[None, :] / lam
    cyc -= np.round(cyc)
    h = np.exp((2j if not reference_formula else -2j) * np.pi * cyc)
    h[:, np.abs(ratio) >= 0.95] = 0
    return h


def f64_rda(raw, p, range_scale=1.0, azimuth_scale=1.0, reference_formula=False):
    """Range compression -> corner turn -> azimuth compression in FP64;
    image is range-bin-major (Nr x Na)."""
    chirp, _ = sar_chirp_reference(p)
    hr = np.conj(np.fft.fft(chirp))
    rc = np.fft.ifft(np.fft.fft(np.asarray(raw, np.complex128), axis=1) * hr, axis=1) * range_scale
    ct = np.ascontiguousarray(rc.T)
    del rc
    ha = sar_azimuth_filter(p, reference_formula)
    return np.fft.ifft(np.fft.fft(ct, axis=1) * ha, axis=1) * azimuth_scale


def sar_targets(p, count, seed):
    """`count` targets at random range / azimuth inside the swath, unit-modulus
    random phase (the synthetic SAR workload)."""
    rng = np.random.default_rng(seed)
    r0 = p["center_range"] + rng.uniform(-1000.0, 1000.0, count)
    span = 0.7 * (p["n_azimuth"] / 2) / p["prf"] * p["velocity"]
    xt = rng.uniform(-span, span, count)
    amp = np.exp(1j * rng.uniform(0, 2 * np.pi, count))
    return list(zip(r0, xt, amp))


def sar_expected_bins(p, r0, xt):
    """Pixel of a focused target (sar_simulator.swift:112-113): range bin,
    azimuth bin."""
    t0 = p["center_range"] * 2.0 / C_LIGHT - (p["n_range"] // 2) / p["sampling_rate"]
    rb = (2.0 * r0 / C_LIGHT - t0) * p["sampling_rate"]
    ab = (xt / p["velocity"] + (p["n_azimuth"] // 2) / p["prf"]) * p["prf"]
    return rb, ab
