"""Python mirror of MetalFFT's convolution / correlation classes over the
libmifft C-ABI.

    Convolver(kernel, fft_size)            Convolve.swift:8-87
    FusedConvolver(kernel, fft_size=4096,  FusedConvolver.swift:17-134
                   precision=Precision.float32)
    Correlator(fft_size)                   Correlate.swift:7-62
    cross_spectral(channels, fft_size,     CrossSpectral.swift:55-173
                   hop_size, window="hann")
    CrossSpectral(fft_size, channels)      (device form of the same, plus
                                            batched STFT and Welch)
    STFT(fftSize, hopSize, window,         STFT.swift:20-99
         sampleRate)
    FFTAnalyzer(size, sampleRate, window)  FFTAnalyzer.swift:66-116
    PSD.welch / PSD.coherence /            PSD.swift:18-134
        PSD.crossSpectral

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
from .api import C64, F32, _count, _dev_ptr, _stream_handle


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
        samples, count = _count(samples, "samples"), _count(count, "count")
        out_len = samples + self.kernelLen - 1
        _lib.check(_lib.load().mifft_convolver_apply_device(
            self._cv, _dev_ptr(signals, 4 * samples * count, F32, "signals"), samples, count,
            _dev_ptr(out, 4 * out_len * count, F32, "out"), _stream_handle(stream)))

    def set_taps_device(self, taps, stream=None) -> None:
        _lib.check(_lib.load().mifft_convolver_set_taps_device(
            self._cv, _dev_ptr(taps, 4 * self.kernelLen, F32, "taps"), _stream_handle(stream)))

    def backward_device(self, signals, grad_out, grad_signal, grad_taps, samples: int,
                        count: int, stream=None) -> None:
        samples, count = _count(samples, "samples"), _count(count, "count")
        out_len = samples + self.kernelLen - 1
        _lib.check(_lib.load().mifft_convolver_backward_device(
            self._cv, _dev_ptr(signals, 4 * samples * count, F32, "signals"), samples, count,
            _dev_ptr(grad_out, 4 * out_len * count, F32, "grad_out"),
            _dev_ptr(grad_signal, 4 * samples * count, F32, "grad_signal"),
            _dev_ptr(grad_taps, 4 * self.kernelLen, F32, "grad_taps"), _stream_handle(stream)))

    def backward_device_noalloc(self, signals, grad_out, grad_signal, grad_taps, samples: int,
                                count: int, stream=None) -> None:
        """backward_device without any device allocation: the tap partials
        live in the library's fixed per-GPU scratch array (calls on one GPU
        must not overlap)."""
        samples, count = _count(samples, "samples"), _count(count, "count")
        out_len = samples + self.kernelLen - 1
        _lib.check(_lib.load().mifft_convolver_backward_device_noalloc(
            self._cv, _dev_ptr(signals, 4 * samples * count, F32, "signals"), samples, count,
            _dev_ptr(grad_out, 4 * out_len * count, F32, "grad_out"),
            _dev_ptr(grad_signal, 4 * samples * count, F32, "grad_signal"),
            _dev_ptr(grad_taps, 4 * self.kernelLen, F32, "grad_taps"), _stream_handle(stream)))

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
        na, nb, count = _count(na, "na"), _count(nb, "nb"), _count(count, "count")
        _lib.check(_lib.load().mifft_correlator_cross_device(
            self._c, _dev_ptr(a, 4 * na * count, F32, "a"), na,
            _dev_ptr(b, 4 * nb * count, F32, "b"), nb, count,
            _dev_ptr(out, 4 * self.fftSize * count, F32, "out"), _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_c", None) and self._c.value:
            _lib.load().mifft_correlator_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _bessel_i0(x: float) -> float:
    """Modified Bessel function I0 by its power series sum ((x/2)^k / k!)^2,
    to 30 terms or a relative term below 1e-12 (Window.swift:99-110)."""
    result = term = 1.0
    h = x / 2.0
    for k in range(1, 31):
        term *= h * h / float(k * k)
        result += term
        if term < 1e-12 * result:
            break
    return result


def window_coefficients(kind, n: int) -> np.ndarray:
    """WindowType.coefficients (Sources/MetalFFT/DSP/Window.swift:14-23) for
    the vDSP windows it calls, from their published definitions (periodic,
    denominator N): hann 0.5 (1 - cos 2 pi n/N) (vDSP_HANN_NORM), hamming
    0.54 - 0.46 cos 2 pi n/N, blackman 0.42 - 0.5 cos 2 pi n/N + 0.08 cos
    4 pi n/N, rectangular 1; flatTop the five-term cosine sum of
    Window.swift:52-66; ("kaiser", beta) I0(beta sqrt(1 - x^2)) / I0(beta),
    x = (n - (N-1)/2) / ((N-1)/2) (Window.swift:69-77; beta 6 for "kaiser")."""
    k = np.arange(n, dtype=np.float64)
    beta = 6.0
    if isinstance(kind, tuple):
        kind, beta = kind[0], float(kind[1])
    if kind == "kaiser":
        half = (n - 1) / 2.0
        x = (k - half) / half
        w = np.array([_bessel_i0(beta * np.sqrt(max(0.0, 1.0 - v * v))) for v in x]) / _bessel_i0(beta)
        return w.astype(np.float32)
    if kind == "flatTop":
        a = (0.21557895, 0.41663158, 0.277263158, 0.083578947, 0.006947368)
        x = k / n
        w = (a[0] - a[1] * np.cos(2 * np.pi * x) + a[2] * np.cos(4 * np.pi * x)
             - a[3] * np.cos(6 * np.pi * x) + a[4] * np.cos(8 * np.pi * x))
        return w.astype(np.float32)
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

    def _opt(self, t, nbytes, kinds, what):
        return _dev_ptr(t, nbytes, kinds, what) if t is not None else None

    def _inputs(self, signals, samples, window):
        samples = _count(samples, "samples")
        n, c = self.fftSize, self.channels
        return (samples, self._opt(signals, 4 * c * samples, F32, "signals"),
                self._opt(window, 4 * n, F32, "window"))

    def compute_device(self, signals, samples: int, hop: int, window, power, cross, coherence,
                       welch_msc=None, stream=None) -> None:
        samples, sig, win = self._inputs(signals, samples, window)
        n, c, p = self.fftSize, self.channels, self.pairs
        _lib.check(_lib.load().mifft_xspec_device(
            self._x, sig, samples, int(hop), win, self._opt(power, 4 * c * n, F32, "power"),
            self._opt(cross, 8 * p * n, C64, "cross"),
            self._opt(coherence, 4 * p * n, F32, "coherence"),
            self._opt(welch_msc, 4 * p * n, F32, "welch_msc"), _stream_handle(stream)))

    def stft_device(self, signals, samples: int, hop: int, window, spectra, power=None,
                    power_db=None, db_alpha: float = 20.0, stream=None) -> None:
        """All frames of all channels windowed and transformed in one launch
        (mifft_stft_device): spectra (C, frames, N) complex64, optional
        power |X|^2 and power_db (C, frames, N) float32."""
        samples, sig, win = self._inputs(signals, samples, window)
        cells = self.channels * self.frames(samples, hop) * self.fftSize
        _lib.check(_lib.load().mifft_stft_device(
            self._x, sig, samples, int(hop), win, self._opt(spectra, 8 * cells, C64, "spectra"),
            self._opt(power, 4 * cells, F32, "power"),
            self._opt(power_db, 4 * cells, F32, "power_db"), float(db_alpha),
            _stream_handle(stream)))

    def welch_device(self, signals, samples: int, hop: int, window, sample_rate: float, psd,
                     stream=None) -> None:
        """PSD.welch of every channel (mifft_welch_device): psd (C, N)."""
        samples, sig, win = self._inputs(signals, samples, window)
        _lib.check(_lib.load().mifft_welch_device(
            self._x, sig, samples, int(hop), win, float(sample_rate),
            self._opt(psd, 4 * self.channels * self.fftSize, F32, "psd"), _stream_handle(stream)))

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
        w = _window(window, N)
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


def _window(window, n: int):
    """Window coefficients as a CUDA tensor (None: rectangular, no multiply)."""
    import torch
    if window is None:
        return None
    spec = isinstance(window, str) or (isinstance(window, tuple) and isinstance(window[0], str))
    w = window_coefficients(window, n) if spec else _f32(window)
    if w.size != n:
        raise FFTError(6, "window length must equal fftSize", expected=n, got=w.size)
    return torch.from_numpy(w).cuda()


class STFTFrame:
    """STFTFrame (Sources/MetalFFT/DSP/STFT.swift:5-15): complex spectrum,
    magnitudes |X|^2, magnitudesDB (Spectrum.toDecibels: 20 log10 as
    vDSP_vdbcon's flag 1 computes, floor -120) and phase atan2(im, re)."""

    def __init__(self, complex_, magnitudes, magnitudes_db=None):
        self.complex = complex_
        self.magnitudes = magnitudes
        self._db = magnitudes_db

    @property
    def magnitudesDB(self) -> np.ndarray:
        if self._db is None:
            self._db = to_decibels(self.magnitudes)
        return self._db

    @property
    def phase(self) -> np.ndarray:
        return np.arctan2(self.complex.imag, self.complex.real).astype(np.float32)


def to_decibels(magnitudes, floor_db: float = -120.0, alpha: float = 20.0) -> np.ndarray:
    """Spectrum.toDecibels (Sources/MetalFFT/DSP/Spectrum.swift:47-55) on the
    host: alpha log10(m) clipped below at floor_db (alpha 20 is what the
    reference's vDSP_vdbcon flag 1 computes)."""
    m = np.asarray(magnitudes, np.float32)
    with np.errstate(divide="ignore"):
        return np.maximum(np.float32(alpha) * np.log10(m), np.float32(floor_db)).astype(np.float32)


class STFT:
    """Short-time Fourier transform (Sources/MetalFFT/DSP/STFT.swift:20-99).
    ``analyze`` windows and transforms every frame of the signal in one
    batched launch (the reference runs one synchronous FFT per frame) and
    returns the same list of STFTFrame."""

    def __init__(self, fftSize: int, hopSize: int, window="hann", sampleRate: float = 44100.0):
        if int(hopSize) < 1:
            raise FFTError(10, "hopSize must be >= 1")
        self.fftSize, self.hopSize, self.sampleRate = int(fftSize), int(hopSize), float(sampleRate)
        self._an = CrossSpectral(self.fftSize, 1)  # unsupportedFFTSize surfaces here
        self._window_spec = window
        self._w = None

    def frameCount(self, sampleCount: int) -> int:
        if sampleCount < self.fftSize:
            return 0
        return (int(sampleCount) - self.fftSize) // self.hopSize + 1

    def binFrequency(self, bin_: int) -> float:
        return bin_ * self.sampleRate / self.fftSize

    def frameTime(self, frameIndex: int) -> float:
        return frameIndex * self.hopSize / self.sampleRate

    def _run(self, signal, want_db: bool):
        import torch
        x = _f32(signal)
        f, n = self.frameCount(x.size), self.fftSize
        if f == 0:
            return None
        if self._w is None:
            self._w = _window(self._window_spec, n)
        dx = torch.from_numpy(x).cuda()
        spec = torch.empty((f, n), dtype=torch.complex64, device="cuda")
        pw = torch.empty((f, n), dtype=torch.float32, device="cuda")
        db = torch.empty((f, n), dtype=torch.float32, device="cuda") if want_db else None
        self._an.stft_device(dx, x.size, self.hopSize, self._w, spec, pw, db)
        torch.cuda.synchronize()
        return spec.cpu().numpy(), pw.cpu().numpy(), None if db is None else db.cpu().numpy()

    def analyze(self, signal) -> list:
        r = self._run(signal, False)
        if r is None:
            return []
        spec, pw, _ = r
        return [STFTFrame(spec[i], pw[i]) for i in range(spec.shape[0])]

    def spectrogram(self, signal) -> np.ndarray:
        """[time][frequency] dB magnitudes (STFT.swift:82-85)."""
        r = self._run(signal, True)
        return np.zeros((0, self.fftSize), np.float32) if r is None else r[2]

    def analyze_device(self, signals, samples: int, spectra, power=None, power_db=None,
                       db_alpha: float = 20.0, stream=None) -> None:
        """Device form over one channel: spectra (frames, N) complex64."""
        if self._w is None:
            self._w = _window(self._window_spec, self.fftSize)
        self._an.stft_device(signals, samples, self.hopSize, self._w, spectra, power, power_db,
                             db_alpha, stream)

    def close(self) -> None:
        self._an.close()


class AnalysisResult:
    """AnalysisResult (Sources/MetalFFT/DSP/FFTAnalyzer.swift:5-61):
    complex spectrum with magnitudes, magnitudesDB, phase, binFrequency and
    rms = sqrt(sum |X|^2) / N.  Pitch and noise heuristics are not part of
    this package."""

    def __init__(self, complex_, magnitudes, sampleRate: float, fftSize: int):
        self.complex, self.magnitudes = complex_, magnitudes
        self.sampleRate, self.fftSize = sampleRate, fftSize

    @property
    def magnitudesDB(self) -> np.ndarray:
        return to_decibels(self.magnitudes)

    @property
    def phase(self) -> np.ndarray:
        return np.arctan2(self.complex.imag, self.complex.real).astype(np.float32)

    def binFrequency(self, bin_: int) -> float:
        return bin_ * self.sampleRate / self.fftSize

    @property
    def rms(self) -> float:
        return float(np.sqrt(np.float32(self.magnitudes.sum(dtype=np.float32)))) / self.fftSize


class FFTAnalyzer:
    """FFTAnalyzer (Sources/MetalFFT/DSP/FFTAnalyzer.swift:66-116): window ->
    FFT of one frame, the one-frame case of the batched STFT launch."""

    def __init__(self, size: int, sampleRate: float, window="hann"):
        self.size, self.sampleRate = int(size), float(sampleRate)
        self._stft = STFT(self.size, 1, window, self.sampleRate)

    def analyze(self, samples, offset: int = 0) -> AnalysisResult:
        x = _f32(samples)
        if offset == 0 and x.size != self.size:
            raise FFTError(6, "samples.count must equal size", expected=self.size, got=x.size)
        if offset < 0 or offset + self.size > x.size:
            raise FFTError(10, "offset + size must be <= samples.count")
        spec, pw, _ = self._stft._run(x[offset:offset + self.size], False)
        return AnalysisResult(spec[0], pw[0], self.sampleRate, self.size)

    def close(self) -> None:
        self._stft.close()


class PSD:
    """PSD (Sources/MetalFFT/DSP/PSD.swift, CrossSpectral.swift): Welch power
    spectral density, magnitude-squared coherence and cross-spectral
    analysis, each over all frames in batched launches."""

    @staticmethod
    def welch(signal, fftSize: int, hopSize: int, sampleRate: float, window="hann") -> np.ndarray:
        """PSD.welch (PSD.swift:18-59): mean |X|^2 / (sum w^2 * sampleRate)."""
        import torch
        if int(hopSize) < 1:
            raise FFTError(10, "hopSize must be >= 1")
        x = _f32(signal)
        an = CrossSpectral(int(fftSize), 1)
        try:
            n = an.fftSize
            w = _window(window, n)
            out = torch.empty(n, dtype=torch.float32, device="cuda")
            dx = torch.from_numpy(x).cuda() if x.size else None
            an.welch_device(dx, x.size, int(hopSize), w, float(sampleRate), out)
            torch.cuda.synchronize()
            return out.cpu().numpy()
        finally:
            an.close()

    @staticmethod
    def coherence(a, b, fftSize: int, hopSize: int, sampleRate: float, window="hann") -> np.ndarray:
        """PSD.coherence (PSD.swift:72-134): |S_ab|^2 / (S_aa S_bb) per bin,
        frames over min(len a, len b), 0 where S_aa S_bb is 0."""
        if int(hopSize) < 1:
            raise FFTError(10, "hopSize must be >= 1")
        a, b = _f32(a), _f32(b)
        m = min(a.size, b.size)
        r = cross_spectral([a[:m], b[:m]], fftSize, hopSize, sampleRate, window)
        return r.welch_msc[0]

    @staticmethod
    def crossSpectral(channels, fftSize: int, hopSize: int, sampleRate: float = 0.0,
                      window="hann") -> CrossSpectralResult:
        return cross_spectral(channels, fftSize, hopSize, sampleRate, window)


__all__ = ["Convolver", "FusedConvolver", "Correlator", "Precision", "FFTError", "CrossSpectral",
           "CrossSpectralResult", "cross_spectral", "window_coefficients", "STFT", "STFTFrame",
           "FFTAnalyzer", "AnalysisResult", "PSD", "to_decibels"]
