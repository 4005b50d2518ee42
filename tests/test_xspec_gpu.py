"""Multi-channel cross-spectral analysis (PSD.crossSpectral,
Sources/MetalFFT/DSP/CrossSpectral.swift:55-173) on the GPU through the
C-ABI, against the FP64 oracle, and the reference's own test cases
(Tests/MetalFFTTests/CrossSpectralTests.swift:68-176)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rnd(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape).astype(np.float32)


@pytest.mark.parametrize("n,hop,c,s", [(1024, 512, 3, 4096), (256, 100, 5, 9000), (64, 64, 1, 640),
                                       (4096, 1024, 4, 40000), (16384, 8192, 2, 50000),
                                       (512, 7, 16, 3000), (64, 1, 2, 100000), (128, 3, 7, 20000)])
def test_cross_spectral_vs_fp64(gpu, oracle, n, hop, c, s):
    x = rnd((c, s), n + c)
    w = gpu.window_coefficients("hann", n)
    r = gpu.cross_spectral(list(x), n, hop, window="hann")
    pw, cr, co, ms = oracle.f64_cross_spectral(x, n, hop, w)
    assert r.power.shape == (c, n) and r.pairCount == c * (c - 1) // 2
    assert oracle.rel_l2(r.power, pw) < 1e-5
    if c > 1:
        assert oracle.rel_l2(r.crossSpectra, cr) < 1e-5
        assert np.max(np.abs(r.coherence - co)) < 1e-4
        assert np.max(np.abs(r.welch_msc - ms)) < 1e-4


def test_identical_channels_unit_coherence(gpu):
    # CrossSpectralTests.swift:68-89
    sig = rnd(1024 * 4, 1)
    r = gpu.cross_spectral([sig, sig, sig], 1024, 512, 44100.0)
    assert r.channels == 3 and r.pairCount == 3
    assert np.allclose(r.coherence[:, 1:], 1.0, atol=1e-4)


def test_single_channel_no_pairs(gpu):
    # CrossSpectralTests.swift:91-103
    r = gpu.cross_spectral([rnd(2048, 2)], 1024, 512)
    assert r.channels == 1 and r.pairCount == 0 and r.power.shape[0] == 1


def test_pair_index_ordering(gpu):
    # CrossSpectralTests.swift:105-121
    r = gpu.cross_spectral([rnd(2048, 3)] * 4, 1024, 512)
    assert [r.pairIndex(i, j) for i, j in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))] == \
        list(range(6))
    assert r.pairCount == 6


def test_partial_coherence_in_range(gpu):
    # CrossSpectralTests.swift:155-176
    s1 = rnd(4096, 4)
    s2 = 0.5 * rnd(4096, 5) + 0.5 * s1
    r = gpu.cross_spectral([s1, s2], 1024, 512)
    assert np.all(r.coherence >= -1e-6) and np.all(r.coherence <= 1 + 1e-6)
    assert np.all(r.welch_msc >= -1e-6) and np.all(r.welch_msc <= 1 + 1e-6)


def test_no_frames_and_preconditions(gpu):
    r = gpu.cross_spectral([rnd(1000, 6), rnd(1000, 7)], 1024, 512)
    assert not r.power.any() and not r.crossSpectra.any() and not r.coherence.any()
    with pytest.raises(gpu.FFTError) as e:
        gpu.cross_spectral([rnd(2048, 8)] * 17, 1024, 512)
    assert e.value.case == "precondition"
    with pytest.raises(gpu.FFTError) as e:
        gpu.CrossSpectral(1000, 2)
    assert e.value.case == "unsupportedFFTSize"
