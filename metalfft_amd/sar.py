"""Python mirror of the SAR range-Doppler processor over the libmifft C-ABI.

Counterpart of the reference's RDA pipelines (src/radar/rda_fused_pipeline.swift,
src/radar/rda_pipeline.swift): range compression, corner turn, azimuth FFT,
range cell migration correction (``rcmc=1``, default), azimuth filter and
IFFT, on complex-FP16 (default) or complex64 device data.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .api import C32H, C64, _dev_ptr, _stream_handle

__all__ = ["RdaParams", "RangeDoppler", "RadarMetrics", "default_targets", "expected_bins",
           "MATCHED", "REFERENCE_FORMULA", "IO_C64", "IO_C32H"]

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

    def _frame(self, t, what):
        """Device address of one n_azimuth x n_range frame in the plan's
        element type (complex FP16 or complex64), size- and type-checked."""
        p = self.params
        c32h = p.io == IO_C32H
        nbytes = (4 if c32h else 8) * p.n_range * p.n_azimuth
        return _dev_ptr(t, nbytes, C32H if c32h else C64, what)

    def process(self, raw, image, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_process_device(
            self._r, self._frame(raw, "raw"), self._frame(image, "image"), _stream_handle(stream)))

    def range_compress(self, raw, out, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_range_compress_device(
            self._r, self._frame(raw, "raw"), self._frame(out, "out"), _stream_handle(stream)))

    def corner_turn(self, src, dst, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_corner_turn_device(
            self._r, self._frame(src, "src"), self._frame(dst, "dst"), _stream_handle(stream)))

    def azimuth_compress(self, src, dst, stream=None) -> None:
        _lib.check(_lib.load().mifft_rda_azimuth_compress_device(
            self._r, self._frame(src, "src"), self._frame(dst, "dst"), _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_r", None) and self._r.value:
            _lib.load().mifft_rda_destroy(self._r)
            self._r = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


C_LIGHT = 299_792_458.0


def default_targets(params: RdaParams) -> list:
    """SARSimulator.defaultTargets (src/radar/sar_simulator.swift:226-244):
    five point targets around the scene centre spaced 50 range resolutions
    (c / 2B) and 50 azimuth pixels (v / PRF); (slant range m, azimuth m,
    amplitude)."""
    r0 = params.center_range
    dr = C_LIGHT / (2.0 * params.bandwidth) * 50
    da = params.velocity / params.prf * 50
    return [(r0, 0.0, 1.0), (r0 + dr, 0.0, 1.0), (r0, da, 1.0), (r0 - dr, -da, 1.0),
            (r0 + 2 * dr, -2 * da, 0.8)]


def expected_bins(params: RdaParams, r0: float, xt: float) -> tuple:
    """Pixel (range bin, azimuth bin) a target at slant range r0 and azimuth
    position xt focuses to (sar_simulator.swift:112-113 geometry)."""
    t0 = params.center_range * 2.0 / C_LIGHT - (params.n_range // 2) / params.sampling_rate
    rb = (2.0 * r0 / C_LIGHT - t0) * params.sampling_rate
    ab = (xt / params.velocity + (params.n_azimuth // 2) / params.prf) * params.prf
    return rb, ab


class RadarMetrics:
    """RadarMetrics (src/radar/radar_metrics.swift:35-267): point-target
    quality of a focused image, on the host.  `image` is this pipeline's
    output layout, [range bin][azimuth bin] complex (the reference's is
    [azimuth][range]; cuts and corners are taken in its sense)."""

    def __init__(self, image, n_range: int, n_azimuth: int):
        import numpy as np
        img = np.asarray(image).reshape(n_range, n_azimuth)
        self.mag = (img.real.astype(np.float32) ** 2 + img.imag.astype(np.float32) ** 2)  # [rg][az]
        self.n_range, self.n_azimuth = n_range, n_azimuth

    def measure_target(self, index: int, expected_range: int, expected_azimuth: int,
                       search_radius: int = 20) -> dict:
        """measureTarget (radar_metrics.swift:55-90)."""
        import numpy as np
        r0 = max(0, expected_range - search_radius)
        r1 = min(self.n_range - 1, expected_range + search_radius)
        a0 = max(0, expected_azimuth - search_radius)
        a1 = min(self.n_azimuth - 1, expected_azimuth + search_radius)
        win = self.mag[r0:r1 + 1, a0:a1 + 1]
        # first maximum in azimuth-major, range-minor order, as findPeak scans
        flat = win.T.reshape(-1)
        k = int(np.argmax(flat))
        pa, pr = a0 + k // win.shape[0], r0 + k % win.shape[0]
        peak = float(self.mag[pr, pa])
        pslr_r, islr_r, res_r = self._cut_metrics(self.mag[:, pa], pr)
        pslr_a, islr_a, res_a = self._cut_metrics(self.mag[pr, :], pa)
        return {"targetIndex": index, "expectedRangeBin": expected_range,
                "expectedAzimuthBin": expected_azimuth, "measuredRangeBin": pr,
                "measuredAzimuthBin": pa, "peakMagnitude": peak, "pslrRange": pslr_r,
                "pslrAzimuth": pslr_a, "islrRange": islr_r, "islrAzimuth": islr_a,
                "snr": self._snr(peak), "resolution3dBRange": res_r,
                "resolution3dBAzimuth": res_a}

    @staticmethod
    def _cut_metrics(cut, peak_bin: int):
        """measureCutMetrics (radar_metrics.swift:131-238): mainlobe = the run
        around the peak above -20 dB; sidelobes within +-min(n/4, 200) bins;
        PSLR and ISLR in dB; -3 dB width by linear interpolation."""
        import math
        n = len(cut)
        pv = float(cut[peak_bin])
        if not pv > 0:
            return 0.0, 0.0, float(n)
        pdb = 10 * math.log10(pv)

        def rel_db(i):
            return 10 * math.log10(cut[i]) - pdb if cut[i] > 0 else -100.0
        left = peak_bin
        for i in range(peak_bin - 1, -1, -1):
            if rel_db(i) < -20.0:
                left = i + 1
                break
            if i == 0:
                left = 0
        right = peak_bin
        for i in range(peak_bin + 1, n):
            if rel_db(i) < -20.0:
                right = i - 1
                break
            if i == n - 1:
                right = n - 1
        main = float(sum(float(cut[i]) for i in range(left, right + 1)))
        half = min(n // 4, 200)
        wl, wr = max(0, peak_bin - half), min(n - 1, peak_bin + half)
        side, mside = 0.0, 0.0
        for i in range(wl, wr + 1):
            if i < left or i > right:
                side += float(cut[i])
                mside = max(mside, float(cut[i]))
        pslr = 10 * math.log10(mside / pv) if mside > 0 else -60.0
        islr = 10 * math.log10(side / main) if side > 0 and main > 0 else -60.0
        t3 = pv * 0.5
        l3 = float(peak_bin)
        for i in range(peak_bin, max(0, peak_bin - half) - 1, -1):
            if cut[i] < t3:
                if i + 1 < n and cut[i + 1] > t3:
                    l3 = i + (t3 - cut[i]) / (cut[i + 1] - cut[i])
                else:
                    l3 = float(i)
                break
        r3 = float(peak_bin)
        for i in range(peak_bin, min(n, peak_bin + half)):
            if cut[i] < t3:
                if i > 0 and cut[i - 1] > t3:
                    r3 = i - (t3 - cut[i]) / (cut[i - 1] - cut[i])
                else:
                    r3 = float(i)
                break
        return pslr, islr, float(max(r3 - l3, 1.0))

    def _snr(self, peak: float) -> float:
        """measureSNR (radar_metrics.swift:240-267): peak power over the mean of
        the two 64 x 64 image corners (first azimuth and range bins, last
        azimuth and range bins)."""
        import math
        c = 64
        noise = float(self.mag[:c, :c].sum() + self.mag[-c:, -c:].sum()) / (2 * c * c)
        if not noise > 0:
            return 60.0
        return 10 * math.log10(peak / noise)
