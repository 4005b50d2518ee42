"""Every *_device entry point captured in a hipGraph (torch.cuda.CUDAGraph)
and replayed several times, with eager GPU work and synchronisations between
the replays: each replay must reproduce the eager result bit for bit (the
entry points issue no memsets and allocate nothing once their scratch has
grown on a first, eager call)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def replay_matches_eager(torch, run, outs, reps=3):
    run()  # eager: scratch grows here, never inside the capture
    torch.cuda.synchronize()
    refs = [o.clone() for o in outs]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    for r in range(reps):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _ = (torch.randn(4096, device="cuda") * r).sum().item()  # eager work between replays
        for o, ref in zip(outs, refs):
            assert torch.equal(o, ref), f"replay {r} differs"
    del g


def test_fft_plans(gpu):
    import torch
    n, b = 4096, 512
    x = torch.randn((b, n), dtype=torch.complex64, device="cuda")
    y, z = torch.empty_like(x), torch.empty_like(x)
    f = gpu.MetalFFT(n)
    mf = gpu.PrecisionFFT(n, gpu.PrecisionFFT.FP16_MFMA_SPLIT)
    w = torch.empty_like(x)

    def run():
        f.execute(x, y, b)
        f.execute(y, z, b, direction=gpu.INVERSE)
        mf.execute(x, w, b)
    replay_matches_eager(torch, run, [y, z, w])


def test_real_input(gpu):
    import torch
    n, b = 2048, 300
    x = torch.randn((b, n), device="cuda")
    y = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
    full = torch.empty((b, n), dtype=torch.complex64, device="cuda")
    back = torch.empty_like(x)
    f = gpu.MetalFFT(n)

    def run():
        f.forward_real(x, y, b, half=True)
        f.forward_real(x, full, b)
        f.inverse_real(y, back, b, half=True)
    replay_matches_eager(torch, run, [y, full, back])


def test_convolver_and_correlator(gpu):
    import torch
    K, S, cnt = 255, 20000, 6
    taps = np.random.default_rng(1).standard_normal(K).astype(np.float32) / K
    cv = gpu.FusedConvolver(taps)
    h = gpu.FusedConvolver(taps, precision=gpu.Precision.float16Storage)
    x = torch.randn((cnt, S), device="cuda")
    gy = torch.randn((cnt, S + K - 1), device="cuda")
    y = torch.empty((cnt, S + K - 1), device="cuda")
    yh = torch.empty_like(y)
    gx = torch.empty_like(x)
    gh = torch.empty(K, device="cuda")
    c = gpu.Correlator(1024)
    a = torch.randn((cnt, 1024), device="cuda")
    r = torch.empty((cnt, 1024), device="cuda")

    def run():
        cv.apply_device(x, y, S, cnt)
        h.apply_device(x, yh, S, cnt)
        cv.backward_device(x, gy, gx, gh, S, cnt)
        c.cross_device(a, 1024, a.flip(1).contiguous(), 1024, cnt, r)
    replay_matches_eager(torch, run, [y, yh, gx, gh, r])


def test_framed_analyses(gpu):
    import torch
    n, hop, ch, s = 1024, 256, 3, 30000
    x = torch.randn((ch, s), device="cuda")
    w = torch.from_numpy(gpu.window_coefficients("hann", n)).cuda()
    an = gpu.CrossSpectral(n, ch)
    f = an.frames(s, hop)
    p = an.pairs
    pw = torch.empty((ch, n), device="cuda")
    cr = torch.empty((p, n), dtype=torch.complex64, device="cuda")
    co = torch.empty((p, n), device="cuda")
    ms = torch.empty((p, n), device="cuda")
    spec = torch.empty((ch, f, n), dtype=torch.complex64, device="cuda")
    db = torch.empty((ch, f, n), device="cuda")
    psd = torch.empty((ch, n), device="cuda")

    def run():
        an.compute_device(x, s, hop, w, pw, cr, co, ms)
        an.stft_device(x, s, hop, w, spec, None, db)
        an.welch_device(x, s, hop, w, 44100.0, psd)
    replay_matches_eager(torch, run, [pw, cr, co, ms, spec, db, psd])


@pytest.mark.parametrize("precision,rcmc", [(0, 1), (2, 1), (0, 0)])
def test_sar_pipeline(gpu, precision, rcmc):
    import torch
    from metalfft_amd.sar import RdaParams, RangeDoppler
    rda = RangeDoppler(RdaParams.defaults(io=1, precision=precision, rcmc=rcmc))
    raw = (torch.randn((4096, 8192), device="cuda") * 0.2).half()
    img = torch.empty_like(raw)

    def run():
        rda.process(raw, img)
    replay_matches_eager(torch, run, [img])


def test_scratch_growth_after_capture(gpu):
    """A graph captured at a small size stays valid after eager calls at a
    larger size have grown the object's scratch (the superseded allocation
    is kept until the object is destroyed): the convolver adjoint's tap
    partials and the analyser's spectra and partial sums both grow here."""
    import torch
    K, S = 255, 20000
    taps = np.random.default_rng(2).standard_normal(K).astype(np.float32) / K
    cv = gpu.FusedConvolver(taps)
    n, hop, ch = 1024, 256, 2
    an = gpu.CrossSpectral(n, ch)
    small, big = 2, 12
    x = torch.randn((big, S), device="cuda")
    gy = torch.randn((big, S + K - 1), device="cuda")
    gx = torch.empty_like(x)
    gh = torch.empty(K, device="cuda")
    sig = torch.randn((ch, 40000), device="cuda")
    pw = torch.empty((ch, n), device="cuda")
    cr = torch.empty((1, n), dtype=torch.complex64, device="cuda")
    co = torch.empty((1, n), device="cuda")

    def run_small():
        cv.backward_device(x, gy, gx, gh, S, small)
        an.compute_device(sig, 8000, hop, None, pw, cr, co)
    run_small()
    torch.cuda.synchronize()
    refs = [t.clone() for t in (gx[:small], gh, pw, cr, co)]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run_small()
    # eager calls at larger sizes grow both objects' scratch
    gx_big = torch.empty_like(x)
    gh_big = torch.empty_like(gh)
    cv.backward_device(x, gy, gx_big, gh_big, S, big)
    an.compute_device(sig, 40000, hop, None, pw, cr, co)
    torch.cuda.synchronize()
    for r in range(3):
        for t in (gx, gh, pw, cr, co):
            t.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _ = torch.randn((1 << 22,), device="cuda").sum().item()  # reuse freed memory, if any
        outs = (gx[:small], gh, pw, cr, co)
        for o, ref in zip(outs, refs):
            assert torch.equal(o, ref), f"replay {r} differs"
    del g
