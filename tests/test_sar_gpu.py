"""SAR range-Doppler focusing (range compression -> corner turn -> azimuth
FFT -> RCMC -> azimuth filter -> IFFT) on the GPU vs the FP64 oracle pipeline
on the same raw data, plus target peak positions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_rda(gpu, oracle, p, targets, io, noise=0.1, seed=7, convention=0, rcmc=1):
    import torch
    from metalfft_amd.sar import RdaParams, RangeDoppler
    raw = oracle.sar_simulate(p, targets)
    rng = np.random.default_rng(seed)
    raw += noise * (rng.standard_normal(raw.shape) + 1j * rng.standard_normal(raw.shape))
    geo = {k: p[k] for k in ("bandwidth", "pulse_duration", "sampling_rate", "prf", "velocity",
                             "carrier_freq", "center_range")}
    prm = RdaParams.defaults(n_range=p["n_range"], n_azimuth=p["n_azimuth"], io=io,
                             filter_convention=convention, rcmc=rcmc, **geo)
    rda = RangeDoppler(prm)
    ri = raw.astype(np.complex64).view(np.float32)
    if io == 1:
        src = torch.from_numpy(ri).cuda().to(torch.float16)
        raw_used = src.float().cpu().numpy().view(np.complex64).astype(np.complex128)
    else:
        src = torch.from_numpy(ri).cuda()
        raw_used = ri.view(np.complex64).astype(np.complex128)
    img = torch.empty((p["n_range"], 2 * p["n_azimuth"]), dtype=src.dtype, device="cuda")
    rda.process(src, img)
    torch.cuda.synchronize()
    got = img.float().cpu().numpy().view(np.complex64)
    ref = oracle.f64_rda(raw_used, p, rda.range_scale, rda.azimuth_scale,
                         reference_formula=bool(convention), rcmc=bool(rcmc))
    return got, ref, rda


@pytest.mark.parametrize("io,bound,rcmc", [(0, 1e-5, 0), (1, 3e-3, 0), (0, 1e-5, 1), (1, 3e-3, 1)])
def test_rda_small_vs_fp64(gpu, oracle, io, bound, rcmc):
    p = oracle.sar_default_params()
    p.update(n_range=2048, n_azimuth=1024)
    tg = oracle.sar_targets(p, 8, 5)
    got, ref, _ = run_rda(gpu, oracle, p, tg, io, rcmc=rcmc)
    err = oracle.rel_l2(got, ref)
    print(f"io={io} rel_l2={err:.3e}")
    assert err < bound
    # peaks: range to a bin; azimuth to 4 bins (the azimuth mainlobe spans
    # ~PRF / Doppler bandwidth ~ 30 bins in this geometry, so its top is flat)
    mag = np.abs(got)
    for r0, xt, _a in tg:
        rb, ab = oracle.sar_expected_bins(p, r0, xt)
        ri, ai = int(round(rb)), int(round(ab))
        w = mag[ri - 3:ri + 4, ai - 8:ai + 9]
        m = np.unravel_index(np.argmax(w), w.shape)
        assert abs(ri - 3 + m[0] - rb) <= 1.0 and abs(ai - 8 + m[1] - ab) <= 4.0


@pytest.mark.parametrize("n_range,n_azimuth", [(4096, 512), (1024, 4096), (4096, 2048)])
@pytest.mark.parametrize("rcmc", [0, 1])
def test_rda_n4096_row_forms(gpu, oracle, n_range, n_azimuth, rcmc):
    """4096-point rows as one workgroup per row (row counts up to the
    resident workgroups: 512 range rows, 1024 azimuth rows) and as the
    persistent form that prefetches the next row (2048 range rows, two per
    workgroup), complex64 data, against the FP64 pipeline."""
    p = oracle.sar_default_params()
    p.update(n_range=n_range, n_azimuth=n_azimuth)
    tg = oracle.sar_targets(p, 6, 11)
    got, ref, _ = run_rda(gpu, oracle, p, tg, 0, rcmc=rcmc)
    err = oracle.rel_l2(got, ref)
    print(f"{n_range}x{n_azimuth} rcmc={rcmc} rel_l2={err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("rcmc", [0, 1])
def test_rda_reference_formula_matches_oracle(gpu, oracle, rcmc):
    p = oracle.sar_default_params()
    p.update(n_range=2048, n_azimuth=512)
    tg = oracle.sar_targets(p, 4, 6)
    got, ref, _ = run_rda(gpu, oracle, p, tg, 0, convention=1, rcmc=rcmc)
    assert oracle.rel_l2(got, ref) < 1e-5


@pytest.mark.parametrize("io,bound", [(0, 1e-5), (1, 3e-3)])
def test_rcmc_large_migration(gpu, oracle, io, bound):
    """L band, 200 m/s, 5 km: up to ~300 range bins of migration across the
    Doppler band; the GPU pipeline with RCMC matches the FP64 one and
    focuses the target on its pixel."""
    p = dict(oracle.sar_default_params())
    p.update(n_range=1024, n_azimuth=2048, carrier_freq=1e9, velocity=200.0, center_range=5000.0)
    tg = [(5000.0, 0.0, 1.0 + 0j), (4980.0, 30.0, np.exp(1j))]
    got, ref, _ = run_rda(gpu, oracle, p, tg, io, noise=0.01)
    assert oracle.rel_l2(got, ref) < bound
    mag = np.abs(got)
    for r0, xt, _a in tg:
        rb, ab = oracle.sar_expected_bins(p, r0, xt)
        ri, ai = int(round(rb)), int(round(ab))
        w = mag[ri - 6:ri + 7, ai - 12:ai + 13]
        m = np.unravel_index(np.argmax(w), w.shape)
        assert abs(ri - 6 + m[0] - rb) <= 1.0 and abs(ai - 12 + m[1] - ab) <= 2.0


def test_rda_steps_compose(gpu, oracle):
    """process == range_compress -> corner_turn -> azimuth_compress."""
    import torch
    from metalfft_amd.sar import RdaParams, RangeDoppler
    rda = RangeDoppler(RdaParams.defaults(n_range=2048, n_azimuth=256, io=1))
    x = (torch.randn((256, 4096), device="cuda") * 0.3).half()
    a = torch.empty_like(x)
    b = torch.empty((2048, 512), dtype=torch.float16, device="cuda")
    rda.range_compress(x, a)
    rda.corner_turn(a, b)
    rda.azimuth_compress(b, b)
    full = torch.empty_like(b)
    rda.process(x, full)
    torch.cuda.synchronize()
    assert torch.equal(full, b)
    # corner turn is an exact transpose
    t = torch.empty((2048, 512), dtype=torch.float16, device="cuda")
    rda.corner_turn(x, t)
    torch.cuda.synchronize()
    xc = x.view(256, 2048, 2).transpose(0, 1).reshape(2048, 512)
    assert torch.equal(t, xc)


def test_sar_fp16_config_4096(gpu, oracle):
    """4096 pulses x 4096 range samples, complex FP16, 32 targets with
    unit-modulus random phase plus N(0, 0.1) noise, vs the FP64 pipeline."""
    p = oracle.sar_default_params()
    tg = oracle.sar_targets(p, 32, 11)
    got, ref, rda = run_rda(gpu, oracle, p, tg, 1)
    err = oracle.rel_l2(got, ref)
    print(f"4096x4096 fp16 rel_l2={err:.3e} scales={rda.range_scale},{rda.azimuth_scale}")
    assert err < 3e-3
    assert np.isfinite(got).all()


@pytest.fixture(scope="module")
def scene_4096(oracle):
    """The 4096 x 4096, 32-target scene as complex FP16 raw data, its FP64
    RDA image (with RCMC) and the targets."""
    import torch
    from metalfft_amd.sar import RdaParams, RangeDoppler
    p = oracle.sar_default_params()
    tg = oracle.sar_targets(p, 32, 11)
    raw = oracle.sar_simulate(p, tg)
    rng = np.random.default_rng(7)
    raw += 0.1 * (rng.standard_normal(raw.shape) + 1j * rng.standard_normal(raw.shape))
    src = torch.from_numpy(raw.astype(np.complex64).view(np.float32)).cuda().to(torch.float16)
    raw_used = src.float().cpu().numpy().view(np.complex64).astype(np.complex128)
    rda = RangeDoppler(RdaParams.defaults(io=1))
    ref = oracle.f64_rda(raw_used, p, rda.range_scale, rda.azimuth_scale, rcmc=True)
    return p, tg, src, ref


@pytest.mark.parametrize("precision,bound", [(1, 6e-3), (2, 4e-3), (3, 5e-3)])
def test_sar_fp16_modes(gpu, oracle, scene_4096, precision, bound):
    """Range and azimuth compression in the reference's FP16 modes (pure,
    storage, mixed; fft_sar_fused_fp16.metal) on the 4096 x 4096 scene: the
    image stays within the mode's bound of the FP64 pipeline, more than FP16
    rounding away from it (the mode really computes in FP16), and every
    target focuses on the same pixel as in FP32."""
    import torch
    from metalfft_amd.sar import RdaParams, RangeDoppler
    p, tg, src, ref = scene_4096
    imgs = {}
    for prec in (0, precision):
        rda = RangeDoppler(RdaParams.defaults(io=1, precision=prec))
        img = torch.empty_like(src)
        rda.process(src, img)
        torch.cuda.synchronize()
        imgs[prec] = img.float().cpu().numpy().view(np.complex64)
    err = oracle.rel_l2(imgs[precision], ref)
    err32 = oracle.rel_l2(imgs[0], ref)
    print(f"precision={precision} rel_l2={err:.3e} (FP32 arithmetic: {err32:.3e})")
    assert np.isfinite(imgs[precision]).all()
    assert err < bound and err > err32
    m16, m32 = np.abs(imgs[precision]), np.abs(imgs[0])
    for r0, xt, _a in tg:
        rb, ab = oracle.sar_expected_bins(p, r0, xt)
        ri, ai = int(round(rb)), int(round(ab))
        w16, w32 = m16[ri - 3:ri + 4, ai - 8:ai + 9], m32[ri - 3:ri + 4, ai - 8:ai + 9]
        assert np.unravel_index(np.argmax(w16), w16.shape)[0] == np.unravel_index(np.argmax(w32), w32.shape)[0]


def test_sar_fp16_mode_without_rcmc_and_errors(gpu, oracle, scene_4096):
    import torch
    from metalfft_amd.sar import RdaParams, RangeDoppler
    p, tg, src, _ = scene_4096
    out = {}
    for prec in (0, 2):
        rda = RangeDoppler(RdaParams.defaults(io=1, precision=prec, rcmc=0))
        img = torch.empty_like(src)
        rda.process(src, img)
        torch.cuda.synchronize()
        out[prec] = img.float().cpu().numpy().view(np.complex64)
    assert oracle.rel_l2(out[2], out[0]) < 1e-2
    with pytest.raises(gpu.FFTError) as e:
        RangeDoppler(RdaParams.defaults(n_range=2048, precision=1))
    assert e.value.case == "kernelNotFound"
    with pytest.raises(gpu.FFTError):
        RangeDoppler(RdaParams.defaults(precision=7))


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_point_targets_pslr_snr(gpu, oracle, precision):
    """The reference's image-quality checks (RadarMetrics) in every precision
    mode: range PSLR below -13 dB for every target
    (precision_comparison.swift:590-598) and SNR above 10 dB (main.swift:213),
    each target on its predicted pixel.  Targets are 400 range and 300
    azimuth bins apart: the reference's default five
    (sar_simulator.swift:226-244) put range neighbours 60 bins apart, inside
    the metric's +-200-bin sidelobe window, where a neighbour's own peak
    counts as a 0 dB sidelobe."""
    import torch
    from metalfft_amd.sar import RadarMetrics, RdaParams, RangeDoppler, expected_bins
    prm = RdaParams.defaults(io=1, precision=precision)
    p = oracle.sar_default_params()
    px = 299_792_458.0 / (2 * prm.sampling_rate)  # range pixel, m
    az = prm.velocity / prm.prf                    # azimuth pixel, m
    r0 = prm.center_range
    tg = [(r0 - 400 * px, -300 * az, 1.0), (r0, 0.0, 1.0), (r0 + 400 * px, 300 * az, 0.8),
          (r0 + 800 * px, -600 * az, 1.0)]
    raw = oracle.sar_simulate(p, [(r, x, complex(a)) for r, x, a in tg])
    rng = np.random.default_rng(3)
    raw += 0.1 * (rng.standard_normal(raw.shape) + 1j * rng.standard_normal(raw.shape))
    src = torch.from_numpy(raw.astype(np.complex64).view(np.float32)).cuda().to(torch.float16)
    img = torch.empty_like(src)
    RangeDoppler(prm).process(src, img)
    torch.cuda.synchronize()
    im = img.float().cpu().numpy().view(np.complex64)
    rm = RadarMetrics(im, prm.n_range, prm.n_azimuth)
    for i, (rt, xt, _a) in enumerate(tg):
        rb, ab = expected_bins(prm, rt, xt)
        m = rm.measure_target(i, int(round(rb)), int(round(ab)))
        print(precision, i, round(m["pslrRange"], 2), round(m["pslrAzimuth"], 2), round(m["snr"], 1))
        assert m["pslrRange"] < -13.0 and m["snr"] > 10.0
        assert abs(m["measuredRangeBin"] - rb) <= 1.0 and abs(m["measuredAzimuthBin"] - ab) <= 4.0
