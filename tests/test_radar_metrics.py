"""RadarMetrics (src/radar/radar_metrics.swift:35-267) on images with known
answers: a sampled sinc^2 point response (PSLR -13.3 dB, -3 dB width
0.886 x the sinc's zero spacing), a known noise floor for SNR, and the
reference's default target layout (sar_simulator.swift:226-244)."""
import numpy as np

from metalfft_amd.sar import RadarMetrics, RdaParams, default_targets, expected_bins


def sinc_image(n, at, q, scale=1.0):
    x = np.arange(n) - at
    img = np.zeros((n, n), np.complex64)
    s = np.sinc(x / q) * scale
    img[:, at] = s
    img[at, :] = s
    return img


def test_sinc_point_response_closed_forms():
    n, at, q = 1024, 500, 8.0
    m = RadarMetrics(sinc_image(n, at, q), n, n).measure_target(0, at + 3, at - 4)
    assert (m["measuredRangeBin"], m["measuredAzimuthBin"]) == (at, at)
    for cut in ("Range", "Azimuth"):
        assert abs(m["pslr" + cut] - (-13.26)) < 0.2  # sampled sidelobe peaks
        assert abs(m["resolution3dB" + cut] - 0.8859 * q) < 0.02 * q
        assert -10.5 < m["islr" + cut] < -9.0


def test_snr_from_corners():
    n = 512
    img = sinc_image(n, 256, 4.0, scale=100.0)
    rng = np.random.default_rng(0)
    noise = 0.5 * (rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64)))
    img[:64, :64] += noise
    img[-64:, -64:] += noise
    m = RadarMetrics(img, n, n).measure_target(0, 256, 256)
    p_noise = 2 * float((np.abs(noise) ** 2).sum()) / (2 * 64 * 64)
    assert abs(m["snr"] - 10 * np.log10(100.0 ** 2 / p_noise)) < 1e-3


def test_default_targets_layout():
    p = RdaParams.defaults()
    tg = default_targets(p)
    assert len(tg) == 5 and tg[0] == (p.center_range, 0.0, 1.0) and tg[4][2] == 0.8
    dr = 299_792_458.0 / (2 * p.bandwidth) * 50
    assert abs(tg[1][0] - tg[0][0] - dr) < 1e-9
    rb, ab = expected_bins(p, p.center_range, 0.0)
    assert abs(rb - p.n_range // 2) < 1e-6 and abs(ab - p.n_azimuth // 2) < 1e-6
