"""Python mirror of the SAR range-Doppler processor over the libmifft C-ABI.

Counterpart of the reference's RDA pipelines (src/radar/rda_fused_pipeline.swift,
src/radar/rda_pipeline.swift): range compression, corner turn, azimuth FFT,
range cell migration correction (``rcmc=1``, default), azimuth filter and
IFFT, on complex-FP16 (default) or complex64 device data.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .api import _dev_ptr, _stream_handle

MATCHED = 0
REFERENCE_FORMULA = 1
IO_C64 = 0
IO_C32H = 1


class RdaParams(ctypes.Structure):
    _fields_ = [("n_range", ctypes.c_int64), ("n_azimuth", ctypes.c_int64),
                ("bandwidth", ctypes.c_double), ("pulse_duration", ctypes.c_double),
                ("sampling_rate", ctypes.c_double), ("prf", ctypes.c_double),
                ("velocity", ctypes.c_double), ("carrier_freq", ctypes.c_double),
                ("center_range", ctypes.c_double), ("io", ctypes.c_int),
                ("filter_convention", ctypes.c_int), ("rcmc", ctypes.c_int),
                ("precision", ctypes.c_int)]

    @classmethod
    def defaults(cls, **overrides) -> "RdaParams":
        p = cls()
        _lib.check(_lib.load().mifft_rda_default_params(ctypes.byref(p)))
        for k, v in overrides.items():
            setattr(p, k, v)
        return p

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class RangeDoppler:
    def __init__(self, params: RdaParams | None = None, **overrides):
        self.params = params if params is not None else RdaParams.defaults(**overrides)
        self._r = ctypes.c_void_p()
        _lib.check(_lib.load().mifft_rda_create(ctypes.byref(self.params), ctypes.byref(self._r)))
        rs, az, ns = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _lib.check(_lib.load().mifft_rda_scales(self._r, ctypes.byref(rs), ctypes.byref(az),
                                                ctypes.byref(ns)))
        self.range_scale, self.azimuth_scale, self.chirp_samples = rs.value, az.value, ns.value

    def process(self, raw, image, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_process_device(self._r, _dev_ptr(raw), _dev_ptr(image),
                                                        _stream_handle(stream)))

    def range_compress(self, raw, out, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_range_compress_device(self._r, _dev_ptr(raw),
                                                               _dev_ptr(out), _stream_handle(stream)))

    def corner_turn(self, src, dst, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_corner_turn_device(self._r, _dev_ptr(src), _dev_ptr(dst),
                                                            _stream_handle(stream)))

    def azimuth_compress(self, src, dst, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_azimuth_compress_device(self._r, _dev_ptr(src),
                                                                 _dev_ptr(dst),
                                                                 _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_r", None) and self._r.value:
            _lib.load().mifft_rda_destroy(self._r)
            self._r = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
