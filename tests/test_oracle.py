"""Pins the CPU oracle (oracle/cpu_fft.c + FP64 numpy) before it is trusted.

Checks, each able to catch a plausible slip (sign, index, dropped term,
transposed operand):
  * numpy FP64 FFT vs the brute-force O(N^2) DFT;
  * closed-form known answers from the reference's tests: impulse -> ones
    (Tests/MetalFFTTests/MetalFFTTests.swift:11-23), DC -> N*delta
    (src/fft_host.swift:527-531), cos at bin 100 (MetalFFTTests.swift:37-49),
    the three-tone signal of src/fft_host.swift:409-421;
  * FP32 restatement vs FP64 at every size, radix-8 vs radix-4, four-step;
  * inverse round trip (Tests/MetalFFTTests/CrossSpectralTests.swift:49-63);
  * overlap-add convolution vs direct FP64 convolution, delta filter
    identity (Tests/MetalFFTTests/FusedConvolverTests.swift:10-21);
  * correlation vs the direct lag sum r[m] = sum_n a[(n+m) mod N] b[n];
  * committed golden vectors (tests/golden/fft_golden.npz).
"""
import numpy as np
import pytest

from conftest import uniform_c64

SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)
# beyond the reference's table: the four-step procedure of MetalFFT.swift:206-277
# with n1 = 2^floor(L/2), n2 = 2^ceil(L/2)
FOURSTEP_SIZES = (32768, 65536, 131072, 262144, 524288, 1 << 20)


@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_numpy_fp64_matches_brute_force(oracle, n):
    x = uniform_c64((n,), 7).astype(np.complex128)
    assert oracle.rel_l2(oracle.f64_fft(x), oracle.dft_brute_f64(x)) < 1e-13
    assert oracle.rel_l2(oracle.dft_brute_f64(x, inverse=True) / n, oracle.f64_ifft(x)) < 1e-13


@pytest.mark.parametrize("n", SIZES)
def test_impulse_gives_ones(oracle, n):
    x = np.zeros(n, np.complex64)
    x[0] = 1
    y = oracle.cpu_fft(x)
    assert np.max(np.abs(y - 1)) < 1e-6


@pytest.mark.parametrize("n", SIZES)
def test_dc_gives_scaled_delta(oracle, n):
    y = oracle.cpu_fft(np.ones(n, np.complex64))
    expect = np.zeros(n)
    expect[0] = n
    assert np.max(np.abs(y - expect)) < 1e-6 * n


def test_cosine_peak_bin(oracle):
    n = 1024
    k = np.arange(n)
    x = np.cos(2 * np.pi * 100 * k / n).astype(np.complex64)
    y = oracle.cpu_fft(x)
    assert int(np.argmax(np.abs(y[: n // 2]))) == 100
    assert abs(y[100] - n / 2) < 1e-3 and abs(y[n - 100] - n / 2) < 1e-3


def test_three_tone_closed_form(oracle):
    # x = cos(2pi 100k/N) + 0.5 cos(2pi 200k/N) + 0.3 sin(2pi 500k/N)
    n = 4096
    t = 2 * np.pi * np.arange(n) / n
    x = (np.cos(100 * t) + 0.5 * np.cos(200 * t) + 0.3 * np.sin(500 * t)).astype(np.complex64)
    expect = np.zeros(n, np.complex128)
    expect[100] = expect[n - 100] = n / 2
    expect[200] = expect[n - 200] = n / 4
    expect[500] = -0.3j * n / 2
    expect[n - 500] = 0.3j * n / 2
    for radix8 in (False, True):
        y = oracle.cpu_fft(x, radix8=radix8)
        assert oracle.rel_l2(y, expect) < 1e-6


def test_sign_error_is_caught(oracle):
    """A transform with the wrong exponent sign (conjugated twiddles) fails
    the three-tone known answer: the check above can see that mistake."""
    n = 4096
    t = 2 * np.pi * np.arange(n) / n
    x = (np.cos(100 * t) + 0.3 * np.sin(500 * t)).astype(np.complex64)
    wrong = np.conj(oracle.cpu_fft(np.conj(x)))  # = sum x e^{+2 pi i nk/N}
    expect = np.zeros(n, np.complex128)
    expect[100] = expect[n - 100] = n / 2
    expect[500] = -0.3j * n / 2
    expect[n - 500] = 0.3j * n / 2
    assert oracle.rel_l2(wrong, expect) > 0.1


@pytest.mark.parametrize("n", SIZES)
def test_fp32_restatement_vs_fp64(oracle, n):
    x = uniform_c64((3, n), n)
    y = oracle.cpu_fft(x)
    assert oracle.rel_l2(y, oracle.f64_fft(x)) < 1e-6
    z = oracle.cpu_fft(y, inverse=True)
    assert oracle.rel_l2(z, x) < 1e-6


def test_radix8_matches_radix4(oracle):
    x = uniform_c64((4, 4096), 3)
    a = oracle.cpu_fft(x, radix8=False)
    b = oracle.cpu_fft(x, radix8=True)
    assert oracle.rel_l2(b, a) < 1e-6
    assert oracle.rel_l2(b, oracle.f64_fft(x)) < 1e-6


@pytest.mark.parametrize("n", [8192, 16384, *FOURSTEP_SIZES])
def test_four_step_index_identity(oracle, n):
    # impulse at position p -> X[k] = W_N^{pk}: pins the four-step's
    # transpose order and twiddle exponent (MetalFFT.swift:206-277)
    p = 3 * n // 7
    x = np.zeros(n, np.complex64)
    x[p] = 1
    k = np.arange(n)
    expect = np.exp(-2j * np.pi * ((p * k) % n) / n)
    assert oracle.max_abs(oracle.cpu_fft(x), expect) < 1e-5


def test_twiddle_table(oracle):
    for n in (64, 4096, 16384):
        tw = oracle.twiddle_table(n)
        exact = np.exp(-2j * np.pi * np.arange(n) / n)
        # rounded once from FP64: within half an FP32 ulp of the exact value
        assert np.max(np.abs(tw.real - exact.real)) <= 2 ** -25
        assert np.max(np.abs(tw.imag - exact.imag)) <= 2 ** -25


def test_golden_vectors(oracle, golden):
    for n in SIZES:
        x, X = golden[f"x_{n}"], golden[f"X_{n}"]
        assert oracle.rel_l2(oracle.f64_fft(x), X) < 1e-14
        assert oracle.rel_l2(oracle.cpu_fft(x), X) < 1e-6
    x = uniform_c64((256, 4096), 0)
    assert np.allclose(np.sum(np.abs(oracle.f64_fft(x)) ** 2, -1), golden["cfg1_energy"], rtol=1e-12)


def test_convolve_matches_direct(oracle):
    rng = np.random.default_rng(5)
    s = rng.standard_normal(9000).astype(np.float32)
    t = rng.standard_normal(129).astype(np.float32)
    y = oracle.cpu_convolve(s, t, 1024)
    direct = np.convolve(s.astype(np.float64), t.astype(np.float64))
    assert y.size == s.size + t.size - 1
    assert oracle.rel_l2(y, direct) < 1e-6
    assert oracle.rel_l2(oracle.f64_convolve(s, t), direct) < 1e-12


def test_convolve_delta_identity(oracle):
    s = np.random.default_rng(6).uniform(-1, 1, 4033).astype(np.float32)
    d = np.zeros(64, np.float32)
    d[0] = 1
    y = oracle.cpu_convolve(s, d, 4096)
    assert oracle.rel_l2(y[: s.size], s) < 1e-6


def test_convolve_rejects_long_kernel(oracle):
    with pytest.raises(ValueError):
        oracle.cpu_convolve(np.ones(10, np.float32), np.ones(64, np.float32), 64)


def test_correlate_matches_lag_sum(oracle):
    rng = np.random.default_rng(8)
    n = 256
    a = rng.standard_normal(300)  # truncated to n (Correlate.swift:54)
    b = rng.standard_normal(100)  # zero-padded
    pa = a[:n]
    pb = np.zeros(n)
    pb[:100] = b
    direct = np.array([sum(pa[(i + m) % n] * pb[i] for i in range(n)) for m in range(n)])
    assert oracle.rel_l2(oracle.f64_correlate(a, b, n), direct) < 1e-12
    assert oracle.rel_l2(oracle.cpu_correlate(a, b, n), direct) < 1e-6


def test_conv_gradients_brute_force(oracle):
    """f64_conv_grads against the direct sums grad_x[n] = sum_k h[k] g[n+k],
    grad_h[k] = sum_s sum_n x[s,n] g[s,n+k], and a central difference."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal((2, 40))
    h = rng.standard_normal(7)
    g = rng.standard_normal((2, 46))
    gx, gh = oracle.f64_conv_grads(x, h, g)
    bx = np.array([[sum(h[k] * g[s, n + k] for k in range(7)) for n in range(40)] for s in range(2)])
    bh = np.array([sum(x[s, n] * g[s, n + k] for s in range(2) for n in range(40)) for k in range(7)])
    assert np.max(np.abs(gx - bx)) < 1e-12 and np.max(np.abs(gh - bh)) < 1e-12
    f = lambda hv: sum(np.dot(np.convolve(x[s], hv), g[s]) for s in range(2))
    e = np.zeros(7); e[3] = 1e-6
    assert abs((f(h + e) - f(h - e)) / 2e-6 - gh[3]) < 1e-6


# ----------------------------------------- FusedConvolver FP16 emulation
def _fp16_case():
    rng = np.random.default_rng(5)
    return (rng.uniform(-1, 1, 4096 + 500).astype(np.float32),
            rng.uniform(-1, 1, 64).astype(np.float32))


def test_fp16_emulation_steps_exact_in_fp64(oracle):
    """With FP64 in place of FP16 / FP32 the emulated reference steps
    (radix-4 passes, twiddles, filter product, conj-FFT-conj, overlap-add)
    give the linear convolution: a wrong index, sign or twiddle fails."""
    x, k = _fp16_case()
    ref = oracle.f64_convolve(x, k)
    for mode in ("pure", "storage", "mixed"):
        y = oracle.fused_convolve_fp16_emulated(x, k, mode, _H=np.float64, _F=np.float64)
        assert oracle.rel_l2(y, ref) < 1e-6, mode


@pytest.mark.parametrize("mode,thr", [("pure", 2e-2), ("storage", 1e-3), ("mixed", 2e-2)])
def test_fp16_emulation_error_magnitude(oracle, mode, thr):
    """FP16 rounding is really applied (error well above FP32's) and stays
    inside the reference's own test thresholds against the FP32 path
    (Tests/MetalFFTTests/FusedConvolverTests.swift:39-49)."""
    x, k = _fp16_case()
    y = oracle.fused_convolve_fp16_emulated(x, k, mode)
    err = oracle.rel_l2(y, oracle.f64_convolve(x, k))
    assert 1e-4 < err < thr
    assert oracle.rel_l2(y, oracle.cpu_convolve(x, k, 4096)) < thr


def test_fp16_emulation_mode_ordering(oracle):
    """FP32 arithmetic with FP16 storage is the most accurate mode."""
    x, k = _fp16_case()
    ref = oracle.f64_convolve(x, k)
    e = {m: oracle.rel_l2(oracle.fused_convolve_fp16_emulated(x, k, m), ref)
         for m in ("pure", "storage", "mixed")}
    assert e["storage"] < e["mixed"] and e["storage"] < e["pure"]


# ------------------------------------------------------------------- RCMC
def test_rcmc_weights_closed_form(oracle):
    """At zero fraction the windowed sinc is the centre tap alone, scaled by
    the Hamming window at (k+3)/7 = 3/7 (rda_kernels.metal:163-179); at
    one half the taps are symmetric about k = 0.5."""
    w = oracle.rcmc_weights(np.array(0.0))
    expect = np.zeros(8)
    expect[3] = 0.54 - 0.46 * np.cos(2 * np.pi * 3 / 7)
    assert np.allclose(w, expect, atol=1e-15)
    h = oracle.rcmc_weights(np.array(0.5))
    win = 0.54 - 0.46 * np.cos(2 * np.pi * np.arange(8) / 7)
    sinc = np.sinc(0.5 - np.arange(-3, 5))
    assert np.allclose(h, sinc * win, atol=1e-15)


def test_rcmc_integer_shift_is_scaled_copy(oracle):
    """An integer shift s reads row r + s: a scaled copy, zero past the edge."""
    rng = np.random.default_rng(3)
    rd = rng.standard_normal((40, 6)) + 1j * rng.standard_normal((40, 6))
    shift = np.array([0.0, 1.0, -2.0, 5.0, -7.0, 3.0])
    out = oracle.rcmc_apply(rd, shift)
    w3 = 0.54 - 0.46 * np.cos(2 * np.pi * 3 / 7)
    for f, s in enumerate(shift.astype(int)):
        for r in range(40):
            src = r + s
            want = w3 * rd[src, f] if 0 <= src < 40 else 0.0
            assert abs(out[r, f] - want) < 1e-12


def test_rcmc_matched_sign_focuses(oracle):
    """Physics pin for the shift formula and its sign: with large range
    migration (L band, 200 m/s, 5 km) a point target focuses to a higher
    peak, on the predicted pixel, with RCMC reading r + shift; the
    reference's literal r - shift defocuses it below no correction."""
    p = dict(oracle.sar_default_params())
    p.update(n_range=1024, n_azimuth=2048, carrier_freq=1e9, velocity=200.0, center_range=5000.0)
    assert np.abs(oracle.sar_rcmc_shift(p)).max() > 100  # bins at the Doppler edge
    raw = oracle.sar_simulate(p, [(5000.0, 0.0, 1.0 + 0j)])
    peaks = {}
    for key, kw in (("none", {}), ("matched", dict(rcmc=True)),
                    ("reference", dict(rcmc=True, reference_formula=True))):
        img = np.abs(oracle.f64_rda(raw, p, **kw))
        peaks[key] = (img.max(), np.unravel_index(np.argmax(img), img.shape))
    rb, ab = oracle.sar_expected_bins(p, 5000.0, 0.0)
    assert peaks["matched"][0] > 1.5 * peaks["none"][0]
    assert peaks["matched"][1] == (int(round(rb)), int(round(ab)))
    assert peaks["reference"][0] < peaks["none"][0]


# -------------------------------------------------------- cross-spectral
def test_cross_spectral_closed_forms(oracle):
    """A cosine with an integer number of cycles per frame (rectangular
    window) has power (A N / 2)^2 at its bin; identical channels give
    cross == power and coherence 1; pairs run (0,1), (0,2), (1,2)."""
    n, hop = 256, 128
    t = np.arange(4 * n)
    a = 0.75 * np.cos(2 * np.pi * 10 * t / n)
    pw, cr, co, ms = oracle.f64_cross_spectral([a, a, 2 * a], n, hop)
    assert abs(pw[0, 10] - (0.75 * n / 2) ** 2) < 1e-6 * pw[0, 10]
    assert np.allclose(cr[0], pw[0], rtol=1e-12, atol=1e-9)
    assert np.allclose(cr[1], 2 * pw[0], rtol=1e-12, atol=1e-9)  # (0, 2)
    assert np.allclose(cr[2], 2 * pw[0], rtol=1e-12, atol=1e-9)  # (1, 2)
    big = pw[0] > 1e-6 * pw[0].max()
    assert np.allclose(co[:, big], 1.0) and np.allclose(ms[:, big], 1.0)


def test_cross_spectral_frames_and_independence(oracle):
    """Frame count (pos + N <= samples); independent white noise: the
    per-frame coherence is 1 (the reference's definition) while the
    coherence of averaged spectra falls to ~1/frames."""
    rng = np.random.default_rng(4)
    n, hop = 128, 64
    x = rng.standard_normal((2, 64 * 64 + 10))
    pw, cr, co, ms = oracle.f64_cross_spectral(x, n, hop)
    frames = (x.shape[1] - n) // hop + 1
    assert frames == 63
    assert np.allclose(co, 1.0)
    assert 0.2 / frames < ms.mean() < 3.0 / frames
    none = oracle.f64_cross_spectral(x[:, :n - 1], n, hop)
    assert all(not np.any(a) for a in none)


# ------------------------------------------------------ STFT / Welch / MSC
def test_stft_frames_parseval_and_brute_force(oracle):
    """Frame count (pos + N <= samples); each frame is the brute-force DFT of
    the windowed slice; Parseval per frame; frame f+1 of x is frame f of
    x shifted by one hop."""
    rng = np.random.default_rng(11)
    n, hop = 64, 24
    x = rng.standard_normal(5 * n + 7)
    w = np.hanning(n + 1)[:n]
    X = oracle.f64_stft(x, n, hop, w)
    assert X.shape == ((x.size - n) // hop + 1, n)
    for f in (0, X.shape[0] - 1):
        seg = x[f * hop:f * hop + n] * w
        assert oracle.rel_l2(X[f], oracle.dft_brute_f64(seg.astype(np.complex128))) < 1e-13
        assert abs((np.abs(X[f]) ** 2).sum() - n * (seg ** 2).sum()) < 1e-9 * n * (seg ** 2).sum()
    assert np.allclose(oracle.f64_stft(x[hop:], n, hop, w)[0], X[1])
    assert oracle.f64_stft(x[:n - 1], n, hop, w).shape == (0, n)


def test_to_decibels_known_values(oracle):
    p = np.array([0.0, 1.0, 100.0, 1e-20, 1e-3])
    assert np.array_equal(oracle.to_decibels(p, 20.0), [-120.0, 0.0, 40.0, -120.0, -60.0])
    assert np.allclose(oracle.to_decibels(p, 10.0), [-120.0, 0.0, 20.0, -120.0, -30.0])


def test_welch_constant_signal_closed_form(oracle):
    """A constant a through the periodic Hann window (sum w = N/2, sum w^2 =
    3N/8): X[0] = aN/2, X[+-1] = -aN/4, all else 0, the same in every frame,
    so psd[0] = 2 a^2 N / (3 fs), psd[1] = psd[N-1] = a^2 N / (6 fs)."""
    from metalfft_amd.dsp import window_coefficients
    n, hop, a, fs = 256, 96, 0.6, 8000.0
    w = window_coefficients("hann", n).astype(np.float64)
    assert abs(w.sum() - n / 2) < 1e-4 and abs((w * w).sum() - 3 * n / 8) < 1e-4
    psd = oracle.f64_welch(np.full(7 * n, a), n, hop, fs, w)
    want = np.zeros(n)
    want[0] = 2 * a * a * n / (3 * fs)
    want[1] = want[-1] = a * a * n / (6 * fs)
    assert np.allclose(psd, want, rtol=1e-5, atol=1e-12)
    assert not np.any(oracle.f64_welch(np.ones(n - 1), n, hop, fs, w))


def test_coherence_closed_forms(oracle):
    """b = 3a: coherence 1 wherever a has power; b = 0: 0 everywhere (the
    denominator guard); independent noise ~ 1/frames; the longer signal is
    cut to the shorter one's length."""
    rng = np.random.default_rng(12)
    n, hop = 128, 64
    a = rng.standard_normal(40 * n)
    c = oracle.f64_coherence(a, 3 * a, n, hop)
    assert np.allclose(c, 1.0)
    assert not np.any(oracle.f64_coherence(a, np.zeros_like(a), n, hop))
    b = rng.standard_normal(a.size + 500)
    c = oracle.f64_coherence(a, b, n, hop)
    frames = (a.size - n) // hop + 1
    assert 0.2 / frames < c.mean() < 3.0 / frames
    assert np.array_equal(c, oracle.f64_coherence(a, b[:a.size], n, hop))
    assert np.allclose(oracle.f64_cross_spectral([a, b[:a.size]], n, hop)[3][0], c)


def test_window_definitions():
    """Periodic cosine windows: sums over a period are closed forms (a0 N);
    Kaiser with beta 0 is rectangular, is symmetric, and its I0 series
    agrees with scipy's Bessel I0."""
    import scipy.special
    from metalfft_amd.dsp import _bessel_i0, window_coefficients
    n = 256
    for kind, a0 in (("hann", 0.5), ("hamming", 0.54), ("blackman", 0.42),
                     ("flatTop", 0.21557895), ("rectangular", 1.0)):
        assert abs(window_coefficients(kind, n).astype(np.float64).sum() - a0 * n) < 1e-4 * n
    assert np.allclose(window_coefficients(("kaiser", 0.0), n), 1.0)
    k = window_coefficients(("kaiser", 8.6), n)
    assert np.allclose(k, k[::-1]) and k.max() <= 1.0 + 1e-7
    for x in (0.0, 0.5, 3.0, 8.6, 14.0):
        assert abs(_bessel_i0(x) - scipy.special.i0(x)) < 1e-11 * scipy.special.i0(x)


@pytest.mark.parametrize("n", FOURSTEP_SIZES)
def test_long_four_step_vs_fp64(oracle, n):
    """Lengths above the reference's table: FP32 four-step vs FP64, DC and
    impulse known answers, and the round trip."""
    x = uniform_c64((2, n), n + 1)
    y = oracle.cpu_fft(x, threads=2)
    assert oracle.rel_l2(y, oracle.f64_fft(x)) < 1e-6
    assert oracle.rel_l2(oracle.cpu_fft(y, inverse=True, threads=2), x) < 1e-6
    d = oracle.cpu_fft(np.ones(n, np.complex64))
    assert abs(d[0] - n) < 1e-6 * n and np.max(np.abs(d[1:])) < 1e-6 * n
    i = np.zeros(n, np.complex64)
    i[0] = 1
    assert np.max(np.abs(oracle.cpu_fft(i) - 1)) < 1e-6


def test_four_step_transposed_split_is_caught(oracle):
    """A four-step with its two sub-lengths swapped in the final transpose
    (writing X[k1 + N1 k2] instead of X[k2 + N2 k1]) is a plausible slip: at
    32768 = 128 x 256 it permutes the bins, and the FP64 comparison above
    would fail on it."""
    n, n1, n2 = 32768, 128, 256
    x = uniform_c64((n,), 5)
    y = oracle.cpu_fft(x)
    k = np.arange(n)
    k2, k1 = k % n2, k // n2
    wrong = np.empty_like(y)
    wrong[k1 + n1 * k2] = y[k2 + n2 * k1]
    assert oracle.rel_l2(wrong, oracle.f64_fft(x)) > 0.5


def test_oracle_rejects_beyond_2_20(oracle):
    with pytest.raises(ValueError):
        oracle.cpu_fft(np.zeros(1 << 21, np.complex64))
