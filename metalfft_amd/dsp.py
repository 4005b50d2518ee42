"""Python mirror of MetalFFT's convolution / correlation classes over the
libmifft C-ABI.

    Convolver(kernel, fft_size)            Convolve.swift:8-87
    FusedConvolver(kernel, fft_size=4096,  FusedConvolver.swift:17-134
                   precision=Precision.float32)
    Correlator(fft_size)                   Correlate.swift:7-62

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


__all__ = ["Convolver", "FusedConvolver", "Correlator", "Precision", "FFTError"]
