"""Batched framing for STFT.analyze, FFTAnalyzer.analyze, PSD.welch and
PSD.coherence (Sources/MetalFFT/DSP/STFT.swift:20-99, FFTAnalyzer.swift:
66-116, PSD.swift:18-134) on the GPU through the C-ABI, against the FP64
oracle.  Every frame of a signal goes through one launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rnd(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape).astype(np.float32)


@pytest.mark.parametrize("n,hop,s,window", [(1024, 256, 10000, "hann"), (64, 1, 700, "hamming"),
                                            (4096, 4096, 4096 * 5 + 3, "blackman"),
                                            (16384, 5000, 60000, "flatTop"),
                                            (512, 300, 4000, ("kaiser", 8.0)),
                                            (128, 128, 128, None)])
def test_stft_vs_fp64(gpu, oracle, n, hop, s, window):
    x = rnd(s, n + hop)
    st = gpu.STFT(n, hop, window=window, sampleRate=48000.0)
    frames = st.analyze(x)
    w = None if window is None else gpu.window_coefficients(window, n)
    X = oracle.f64_stft(x, n, hop, w)
    assert len(frames) == st.frameCount(s) == X.shape[0]
    spec = np.stack([f.complex for f in frames])
    mags = np.stack([f.magnitudes for f in frames])
    assert oracle.rel_l2(spec, X) < 1e-6
    assert oracle.rel_l2(mags, np.abs(X) ** 2) < 2e-6
    db = st.spectrogram(x)
    want = oracle.to_decibels(np.abs(X) ** 2, 20.0)
    big = np.abs(X) ** 2 > 1e-6 * (np.abs(X) ** 2).max()
    assert np.max(np.abs(db - want)[big]) < 1e-3
    assert np.max(np.abs(db - np.stack([f.magnitudesDB for f in frames]))) < 1e-4
    st.close()


def test_stft_decibel_floor_and_factor(gpu, oracle):
    """A zero signal gives the -120 dB floor in every bin; a cosine with 8
    cycles per frame (rectangular window) has |X|^2 = (N/2)^2 at bins 8 and
    N-8: 20 log10 or 10 log10 of that per db_alpha."""
    import torch
    n = 256
    an = gpu.CrossSpectral(n, 2)
    t = np.arange(2 * n)
    x = np.stack([np.zeros(2 * n), np.cos(2 * np.pi * 8 * t / n)]).astype(np.float32)
    dx = torch.from_numpy(x).cuda()
    f = an.frames(2 * n, n)
    spec = torch.empty((2, f, n), dtype=torch.complex64, device="cuda")
    for alpha in (20.0, 10.0):
        db = torch.empty((2, f, n), dtype=torch.float32, device="cuda")
        an.stft_device(dx, 2 * n, n, None, spec, None, db, alpha)
        torch.cuda.synchronize()
        d = db.cpu().numpy()
        assert np.all(d[0] == -120.0)
        want = alpha * np.log10((n / 2) ** 2)
        assert np.max(np.abs(d[1][:, [8, n - 8]] - want)) < 1e-3
    an.close()


def test_multichannel_stft_device(gpu, oracle):
    import torch
    n, hop, c, s = 2048, 512, 5, 20000
    x = rnd((c, s), 3)
    w = gpu.window_coefficients("hann", n)
    an = gpu.CrossSpectral(n, c)
    f = an.frames(s, hop)
    spec = torch.empty((c, f, n), dtype=torch.complex64, device="cuda")
    pw = torch.empty((c, f, n), dtype=torch.float32, device="cuda")
    an.stft_device(torch.from_numpy(x).cuda(), s, hop, torch.from_numpy(w).cuda(), spec, pw)
    torch.cuda.synchronize()
    for ch in range(c):
        X = oracle.f64_stft(x[ch], n, hop, w)
        assert oracle.rel_l2(spec[ch].cpu().numpy(), X) < 1e-6
        assert oracle.rel_l2(pw[ch].cpu().numpy(), np.abs(X) ** 2) < 2e-6
    an.close()


def test_fft_analyzer(gpu, oracle):
    n = 1024
    x = rnd(3000, 9)
    fa = gpu.FFTAnalyzer(n, 44100.0)
    w = gpu.window_coefficients("hann", n)
    r = fa.analyze(x, offset=700)
    X = np.fft.fft(x[700:700 + n].astype(np.float64) * w)
    assert oracle.rel_l2(r.complex, X) < 1e-6
    assert abs(r.rms - np.sqrt((np.abs(X) ** 2).sum()) / n) < 1e-5 * r.rms
    assert abs(r.binFrequency(10) - 10 * 44100.0 / n) < 1e-9
    with pytest.raises(gpu.FFTError) as e:
        fa.analyze(x)
    assert e.value.case == "invalidInputSize"
    fa.close()


@pytest.mark.parametrize("n,hop,s,fs", [(1024, 512, 44100, 44100.0), (64, 10, 5000, 1000.0),
                                        (8192, 2048, 100000, 96000.0), (64, 1, 200000, 8000.0),
                                        (16384, 3000, 16384, 48000.0), (256, 7, 300000, 44100.0)])
def test_welch_vs_fp64(gpu, oracle, n, hop, s, fs):
    x = rnd(s, n)
    got = gpu.PSD.welch(x, n, hop, fs)
    want = oracle.f64_welch(x, n, hop, fs, gpu.window_coefficients("hann", n))
    assert oracle.rel_l2(got, want) < 1e-5


def test_welch_constant_closed_form_and_no_frames(gpu):
    n, fs = 256, 8000.0
    got = gpu.PSD.welch(np.full(7 * n, 0.6, np.float32), n, 96, fs)
    assert abs(got[0] - 2 * 0.36 * n / (3 * fs)) < 1e-5 * got[0]
    assert abs(got[1] - 0.36 * n / (6 * fs)) < 1e-5 * got[1]
    assert np.max(np.abs(got[2:-1])) < 1e-6 * got[0]
    assert not gpu.PSD.welch(np.ones(n - 1, np.float32), n, 96, fs).any()
    with pytest.raises(gpu.FFTError):
        gpu.PSD.welch(np.ones(n, np.float32), n, 96, 0.0)


def test_welch_multichannel_matches_power(gpu, oracle):
    """Channel c of mifft_welch_device = cross-spectral power[c] / (sum w^2 fs)
    (CrossSpectralTests.swift:123-153 checks both are positive)."""
    import torch
    n, hop, c, s, fs = 1024, 256, 3, 12000, 44100.0
    x = rnd((c, s), 21)
    w = gpu.window_coefficients("hann", n)
    an = gpu.CrossSpectral(n, c)
    psd = torch.empty((c, n), dtype=torch.float32, device="cuda")
    an.welch_device(torch.from_numpy(x).cuda(), s, hop, torch.from_numpy(w).cuda(), fs, psd)
    torch.cuda.synchronize()
    got = psd.cpu().numpy()
    for ch in range(c):
        assert oracle.rel_l2(got[ch], oracle.f64_welch(x[ch], n, hop, fs, w)) < 1e-5
    r = gpu.cross_spectral(list(x), n, hop)
    assert oracle.rel_l2(got, r.power / (float((w.astype(np.float64) ** 2).sum()) * fs)) < 1e-5
    an.close()


def test_coherence_vs_fp64(gpu, oracle):
    n, hop = 512, 128
    a = rnd(20000, 31)
    b = 0.5 * rnd(20500, 32) + 0.5 * np.concatenate([a, np.zeros(500, np.float32)])
    got = gpu.PSD.coherence(a, b, n, hop, 44100.0)
    want = oracle.f64_coherence(a, b, n, hop, gpu.window_coefficients("hann", n))
    assert np.max(np.abs(got - want)) < 1e-4
    assert not gpu.PSD.coherence(a, np.zeros_like(a), n, hop, 44100.0).any()
    assert np.allclose(gpu.PSD.coherence(a, 3 * a, n, hop, 44100.0), 1.0, atol=1e-4)


def test_host_entry_points_match_device(gpu, oracle):
    """mifft_stft_host / mifft_welch_host / mifft_xspec_host (blocking, host
    buffers) give the device forms' results."""
    import ctypes
    from metalfft_amd import _lib
    L = _lib.load()
    n, hop, c, s = 512, 200, 3, 6000
    x = rnd((c, s), 41)
    w = gpu.window_coefficients("hann", n)
    fp = lambda a: a.ctypes.data_as(_lib.fptr)  # noqa: E731
    an = gpu.CrossSpectral(n, c)
    f = an.frames(s, hop)
    spec = np.empty((c, f, n), np.complex64)
    pw = np.empty((c, f, n), np.float32)
    _lib.check(L.mifft_stft_host(an._x, fp(x), s, hop, fp(w), fp(spec.view(np.float32)), fp(pw), None,
                                 ctypes.c_float(20.0)))
    for ch in range(c):
        X = oracle.f64_stft(x[ch], n, hop, w)
        assert oracle.rel_l2(spec[ch], X) < 1e-6
        assert oracle.rel_l2(pw[ch], np.abs(X) ** 2) < 2e-6
    psd = np.empty((c, n), np.float32)
    _lib.check(L.mifft_welch_host(an._x, fp(x), s, hop, fp(w), ctypes.c_double(1000.0), fp(psd)))
    for ch in range(c):
        assert oracle.rel_l2(psd[ch], oracle.f64_welch(x[ch], n, hop, 1000.0, w)) < 1e-5
    p = c * (c - 1) // 2
    power = np.empty((c, n), np.float32)
    cross = np.empty((p, n), np.complex64)
    coh = np.empty((p, n), np.float32)
    msc = np.empty((p, n), np.float32)
    _lib.check(L.mifft_xspec_host(an._x, fp(x), s, hop, fp(w), fp(power), fp(cross.view(np.float32)),
                                  fp(coh), fp(msc)))
    pw64, cr64, co64, ms64 = oracle.f64_cross_spectral(x, n, hop, w)
    assert oracle.rel_l2(power, pw64) < 1e-5 and oracle.rel_l2(cross, cr64) < 1e-5
    assert np.max(np.abs(coh - co64)) < 1e-4 and np.max(np.abs(msc - ms64)) < 1e-4
    an.close()
