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


# ------------------------------------------- FusedConvolver FP16 emulation
# The reference's FP16 fused convolution (Sources/MetalFFT/Resources/
# fft_fused_convolve_fp16.metal:160-498) restated in numpy, step by step:
# 6 radix-4 Stockham passes (tid < 1024 reads buf[tid + j*1024]; pass p > 0
# twiddles W^{pos*1024/stride} from FP32 sincos with w2 = w1*w1, w3 = w2*w1;
# writes grp*4*stride + pos + j*stride), threadgroup storage in FP16 after
# every pass, filter product, conj-FFT-conj inverse, 1/N scale; host
# overlap-add as FusedConvolver.apply (FusedConvolver.swift:91-133).  numpy
# float16 arithmetic rounds every operation to FP16 (the Apple GPU may fuse
# multiply-adds), so this pins the error magnitude of each mode, not bits.
_H = np.float16
_F = np.float32


def _cmul(ar, ai, br, bi):
    return ar * br - ai * bi, ar * bi + ai * br


def _fp16_pass(xr, xi, p, mode, _H=_H, _F=_F):
    """Pass p of the N = 4096 radix-4 transform on (B, 4096) arrays holding
    the threadgroup buffer (FP16 for every mode); returns the new buffer."""
    n, q = 4096, 1024
    stride = 4 ** p
    tid = np.arange(q)
    pos, grp = tid & (stride - 1), tid // stride
    ar = [xr[:, j * q:(j + 1) * q] for j in range(4)]
    ai = [xi[:, j * q:(j + 1) * q] for j in range(4)]
    if mode != "pure":
        ar = [a.astype(_F) for a in ar]
        ai = [a.astype(_F) for a in ai]
    if p > 0:
        ang = _F(-2.0 * np.pi / n) * (pos * (q // stride)).astype(_F)
        c1, s1 = np.cos(ang).astype(_F), np.sin(ang).astype(_F)
        if mode == "storage":
            w1 = (c1, s1)
        else:
            w1 = (c1.astype(_H), s1.astype(_H))
        w2 = _cmul(*w1, *w1)
        w3 = _cmul(*w2, *w1)
        for j, w in ((1, w1), (2, w2), (3, w3)):
            if mode == "mixed":
                r, i = _cmul(ar[j].astype(_H), ai[j].astype(_H), *w)
                ar[j], ai[j] = r.astype(_F), i.astype(_F)
            else:
                ar[j], ai[j] = _cmul(ar[j], ai[j], *w)
    t0 = (ar[0] + ar[2], ai[0] + ai[2])
    t1 = (ar[1] + ar[3], ai[1] + ai[3])
    t2 = (ar[0] - ar[2], ai[0] - ai[2])
    t3 = (ai[1] - ai[3], -(ar[1] - ar[3]))  # (x1 - x3) * -i
    y = [(t0[0] + t1[0], t0[1] + t1[1]), (t2[0] + t3[0], t2[1] + t3[1]),
         (t0[0] - t1[0], t0[1] - t1[1]), (t2[0] - t3[0], t2[1] - t3[1])]
    wr = grp * 4 * stride + pos
    outr = np.empty((xr.shape[0], n), dtype=y[0][0].dtype)
    outi = np.empty_like(outr)
    for j in range(4):
        outr[:, wr + j * stride] = y[j][0]
        outi[:, wr + j * stride] = y[j][1]
    return outr, outi


def fused_convolve_fp16_emulated(signal: np.ndarray, taps: np.ndarray, mode: str,
                                 _H=_H, _F=_F) -> np.ndarray:
    """FusedConvolver(kernel: taps, fftSize: 4096, precision: .float16<mode>)
    .apply(to: signal), mode in {"pure", "storage", "mixed"}.  (_H, _F
    replace the FP16 / FP32 types; with float64 for both the steps must give
    the exact linear convolution, which the tests check.)"""
    n = 4096
    x = np.asarray(signal, np.float32)
    k = np.asarray(taps, np.float32)
    hop = n - k.size + 1
    spec = np.fft.fft(np.pad(k.astype(np.float64), (0, n - k.size))).astype(np.complex64)
    nblk = (x.size + hop - 1) // hop
    blocks = np.zeros((nblk, n), np.float32)
    for b in range(nblk):
        seg = x[b * hop:(b + 1) * hop]
        blocks[b, :seg.size] = seg
    # input load: half2(src) for pure, FP32 otherwise; pass 0 reads it
    if mode == "pure":
        br, bi = blocks.astype(_H), np.zeros_like(blocks, dtype=_H)
    else:
        br, bi = blocks, np.zeros_like(blocks)
    for p in range(6):
        br, bi = _fp16_pass(br, bi, p, mode, _H, _F)
        br, bi = br.astype(_H), bi.astype(_H)  # threadgroup storage
    fr, fi = spec.real.astype(_F), spec.imag.astype(_F)
    if mode == "storage":
        r, i = _cmul(br.astype(_F), bi.astype(_F), fr, fi)
    else:
        r, i = _cmul(br, bi, fr.astype(_H), fi.astype(_H))
    br, bi = r.astype(_H), (-i).astype(_H)  # stored, then conjugated in place
    for p in range(6):
        br, bi = _fp16_pass(br, bi, p, mode, _H, _F)
        if p < 5:
            br, bi = br.astype(_H), bi.astype(_H)
    if mode == "pure":
        yb = (br.astype(_H) * _H(1.0 / n)).astype(np.float64)
    else:
        yb = (br.astype(_F) * _F(1.0 / n)).astype(np.float64)
    out = np.zeros(x.size + k.size - 1)
    for b in range(nblk):
        lo = b * hop
        hi = min(lo + n, out.size)
        out[lo:hi] += yb[b, :hi - lo]
    return out


# ------------------------------------------------ cross-spectral (FP64)
def f64_cross_spectral(channels, fft_size, hop, window=None):
    """PSD.crossSpectral (Sources/MetalFFT/DSP/CrossSpectral.swift:55-173) in
    FP64: frames at pos = 0, hop, ... while pos + N <= samples, each channel
    windowed and transformed; per frame |X_c|^2, S_ij = X_i conj(X_j) over the
    upper-triangle pairs (0,1), (0,2), ..., and |S_ij|^2 / max(P_i P_j,
    1e-30) (fft_cross_spectral.metal:29-64), all averaged over frames; plus
    the coherence of the averaged spectra.  Returns (power, cross, coherence,
    welch_msc); zeros when there is no frame."""
    x = np.asarray(channels, np.float64)
    c, s = x.shape
    n = int(fft_size)
    w = np.ones(n) if window is None else np.asarray(window, np.float64)
    pairs = [(i, j) for i in range(c) for j in range(i + 1, c)]
    starts = list(range(0, s - n + 1, hop)) if s >= n else []
    power = np.zeros((c, n))
    cross = np.zeros((len(pairs), n), np.complex128)
    coh = np.zeros((len(pairs), n))
    for pos in starts:
        X = np.fft.fft(x[:, pos:pos + n] * w, axis=1)
        P = np.abs(X) ** 2
        power += P
        for q, (i, j) in enumerate(pairs):
            S = X[i] * np.conj(X[j])
            cross[q] += S
            coh[q] += np.abs(S) ** 2 / np.maximum(P[i] * P[j], 1e-30)
    f = max(1, len(starts))
    power, cross, coh = power / f, cross / f, coh / f
    msc = np.zeros_like(coh)
    for q, (i, j) in enumerate(pairs):
        msc[q] = _msc(cross[q], power[i], power[j])
    if not starts:
        msc[:] = 0
    return power, cross, coh, msc


def _msc(sxy, sxx, syy):
    """|S_xy|^2 / (S_xx S_yy), 0 where the denominator is not > 0
    (PSD.coherence, Sources/MetalFFT/DSP/PSD.swift:127-133)."""
    den = sxx * syy
    out = np.zeros(np.shape(den))
    ok = den > 0
    out[ok] = np.abs(sxy[ok]) ** 2 / den[ok]
    return out


# ------------------------------------------------ STFT / Welch (FP64)
def f64_stft(signal, fft_size, hop, window=None):
    """STFT.analyze (Sources/MetalFFT/DSP/STFT.swift:71-99): frames at
    pos = 0, hop, ... while pos + N <= samples, each windowed and
    transformed.  Returns complex spectra [frames][N] (empty when there is
    no frame)."""
    x = np.asarray(signal, np.float64).reshape(-1)
    n = int(fft_size)
    w = np.ones(n) if window is None else np.asarray(window, np.float64)
    starts = list(range(0, x.size - n + 1, hop)) if x.size >= n else []
    if not starts:
        return np.zeros((0, n), np.complex128)
    return np.fft.fft(np.stack([x[p:p + n] * w for p in starts]), axis=1)


def to_decibels(power, alpha=20.0, floor_db=-120.0):
    """Spectrum.toDecibels (Sources/MetalFFT/DSP/Spectrum.swift:47-55):
    alpha * log10(power) (vDSP_vdbcon: alpha 20 for its flag 1, 10 for flag
    0), clipped below at floor_db; zero power gives the floor."""
    p = np.asarray(power, np.float64)
    with np.errstate(divide="ignore"):
        db = alpha * np.log10(p)
    return np.maximum(db, floor_db)


def f64_welch(signal, fft_size, hop, sample_rate, window=None):
    """PSD.welch (Sources/MetalFFT/DSP/PSD.swift:18-59): sum over frames of
    |X|^2 / (frames * sum(w^2) * sample_rate); zeros when there is no
    frame."""
    X = f64_stft(signal, fft_size, hop, window)
    n = int(fft_size)
    if X.shape[0] == 0:
        return np.zeros(n)
    w = np.ones(n) if window is None else np.asarray(window, np.float64)
    return (np.abs(X) ** 2).sum(axis=0) / (X.shape[0] * float((w * w).sum()) * sample_rate)


def f64_coherence(a, b, fft_size, hop, window=None):
    """PSD.coherence (Sources/MetalFFT/DSP/PSD.swift:72-134): frames over
    min(len a, len b); |sum A conj B|^2 / (sum |A|^2 sum |B|^2), 0 where the
    denominator is not > 0 or there is no frame."""
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    m = min(a.size, b.size)
    A = f64_stft(a[:m], fft_size, hop, window)
    B = f64_stft(b[:m], fft_size, hop, window)
    if A.shape[0] == 0:
        return np.zeros(int(fft_size))
    return _msc((A * np.conj(B)).sum(axis=0), (np.abs(A) ** 2).sum(axis=0),
                (np.abs(B) ** 2).sum(axis=0))


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


def sar_rcmc_shift(p, reference_formula=False):
    """Signed source offset, in range bins, per azimuth-frequency bin for the
    RCMC of src/radar/rda_kernels.metal:103-185, in FP64: the migration
    dR(f_a) = R0 (1 / sqrt(1 - ratio^2) - 1) (R0 the centre range, ratio =
    lambda f_a / 2V with f_a = fa_norm * PRF in FFT order, dR = 0 where
    |ratio| >= 0.95) over the range pixel c / 2 fs.  The output bin r reads
    r + shift: energy migrated to a longer range is pulled back (+dR); with
    reference_formula the kernel's literal srcIdx = rgIdx - shift (-dR)."""
    na = p["n_azimuth"]
    lam = C_LIGHT / p["carrier_freq"]
    pixel = C_LIGHT / (2.0 * p["sampling_rate"])
    k = np.arange(na)
    fa = np.where(k < na // 2, k, k - na) / na * p["prf"]
    ratio = lam * fa / (2.0 * p["velocity"])
    dr = np.where(np.abs(ratio) < 0.95,
                  p["center_range"] * (1.0 / np.sqrt(np.maximum(1e-300, 1.0 - ratio ** 2)) - 1.0),
                  0.0)
    return (-1.0 if reference_formula else 1.0) * dr / pixel


def rcmc_weights(frac):
    """The kernel's 8 interpolation weights for taps k = -3 .. 4 at
    fractional offset frac (rda_kernels.metal:163-179): sinc(frac - k) (1
    within 1e-6 of 0) times the Hamming window 0.54 - 0.46 cos(2 pi (k+3)/7)."""
    frac = np.asarray(frac, np.float64)
    out = []
    for k in range(-3, 5):
        x = frac - k
        sinc = np.where(np.abs(x) < 1e-6, 1.0, np.sin(np.pi * x) / np.where(x == 0, 1.0, np.pi * x))
        out.append(sinc * (0.54 - 0.46 * np.cos(2.0 * np.pi * (k + 3) / 7.0)))
    return np.stack(out)  # (8, ...)


def f64_rcmc(rd, p, reference_formula=False):
    """RCMC on range-Doppler data laid out range-bin-major (Nr x Na, the
    layout after the corner turn and azimuth FFT): out[r, f] = sum over
    k = -3..4 of rd[r + c(f) + k, f] w_k(frac(f)), shift(f) = c(f) + frac(f)
    (the same for every range bin), taps outside [0, Nr) dropped
    (rda_kernels.metal:150-180)."""
    return rcmc_apply(rd, sar_rcmc_shift(p, reference_formula))


def rcmc_apply(rd, shift):
    """The interpolation of f64_rcmc for an explicit per-column shift."""
    rd = np.asarray(rd, np.complex128)
    nr, na = rd.shape
    c = np.floor(shift)
    w = rcmc_weights(shift - c)  # (8, Na)
    c = c.astype(np.int64)
    out = np.zeros_like(rd)
    rows = np.arange(nr)[:, None]
    cols = np.arange(na)[None, :]
    for i, k in enumerate(range(-3, 5)):
        idx = rows + c[None, :] + k
        ok = (idx >= 0) & (idx < nr)
        out += np.where(ok, rd[np.clip(idx, 0, nr - 1), cols], 0) * w[i][None, :]
    return out


def f64_rda(raw, p, range_scale=1.0, azimuth_scale=1.0, reference_formula=False, rcmc=False):
    """Range compression -> corner turn -> azimuth FFT -> (RCMC) -> azimuth
    filter -> azimuth IFFT in FP64; image is range-bin-major (Nr x Na)."""
    chirp, _ = sar_chirp_reference(p)
    hr = np.conj(np.fft.fft(chirp))
    rc = np.fft.ifft(np.fft.fft(np.asarray(raw, np.complex128), axis=1) * hr, axis=1) * range_scale
    ct = np.ascontiguousarray(rc.T)
    del rc
    rd = np.fft.fft(ct, axis=1)
    del ct
    if rcmc:
        rd = f64_rcmc(rd, p, reference_formula)
    ha = sar_azimuth_filter(p, reference_formula)
    return np.fft.ifft(rd * ha, axis=1) * azimuth_scale


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
