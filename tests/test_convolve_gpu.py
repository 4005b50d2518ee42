"""Fused FFT-convolve, its adjoint and the correlator on the GPU, through the
C-ABI, against the oracle's overlap-add restatement (Convolve.swift:36-76),
FP64 convolution / gradients, central finite differences and the adjoint
identity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rnd(shape, seed, scale=1.0):
    return (np.random.default_rng(seed).standard_normal(shape) * scale).astype(np.float32)


@pytest.mark.parametrize("n,k,s", [(64, 1, 1000), (64, 63, 333), (128, 17, 5000), (256, 100, 2049),
                                   (512, 129, 4000), (1024, 129, 9000), (2048, 1000, 20000),
                                   (4096, 1023, 65536), (8192, 3000, 50000), (16384, 16383, 40000),
                                   (4096, 64, 1), (4096, 64, 4033)])
def test_convolver_matches_oracle_and_fp64(gpu, oracle, n, k, s):
    sig = rnd(s, n + k)
    taps = rnd(k, s + 1, 1 / np.sqrt(k))
    cv = gpu.Convolver(taps, n)
    assert cv.blockSize == n and cv.hopSize == n - k + 1 and cv.kernelLen == k
    y = cv.apply(sig)
    assert y.size == s + k - 1
    ref64 = oracle.f64_convolve(sig, taps)
    cpu = oracle.cpu_convolve(sig, taps, n)
    assert oracle.rel_l2(y, ref64) < 1e-6
    # 1e-6 to the CPU path; where that path is itself further than 1e-6 from
    # FP64 (hop = 2 at N = 16384: 20,000 overlapping blocks) its own error is
    # the slack
    assert oracle.rel_l2(y, cpu) < 1e-6 + oracle.rel_l2(cpu, ref64)


def test_delta_filter_is_identity(gpu, oracle):
    # Tests/MetalFFTTests/FusedConvolverTests.swift:10-21
    d = np.zeros(64, np.float32)
    d[0] = 1
    conv = gpu.FusedConvolver(d, 4096)
    sig = np.random.default_rng(1).uniform(-1, 1, conv.hopSize).astype(np.float32)
    out = conv.apply(sig)
    assert oracle.rel_l2(out[: sig.size], sig) < 1e-6


def test_fused_matches_convolver(gpu, oracle):
    # FusedConvolverTests.swift:23-35
    taps = np.random.default_rng(2).uniform(-1, 1, 128).astype(np.float32)
    sig = np.random.default_rng(3).uniform(-1, 1, 8192).astype(np.float32)
    a = gpu.FusedConvolver(taps, 4096).apply(sig)
    b = gpu.Convolver(taps, 4096).apply(sig)
    assert oracle.rel_l2(a, b) < 1e-6


def test_empty_signal(gpu):
    cv = gpu.Convolver(np.ones(5, np.float32), 64)
    out = cv.apply(np.zeros(0, np.float32))
    assert out.size == 4 and not out.any()


def test_convolver_preconditions(gpu):
    with pytest.raises(gpu.FFTError) as ei:
        gpu.Convolver(np.ones(64, np.float32), 64)  # Convolve.swift:20-21
    assert ei.value.case == "precondition"
    with pytest.raises(gpu.FFTError) as ei:
        gpu.FusedConvolver(np.ones(8, np.float32), 2048)  # FusedConvolver.swift:56-57
    assert ei.value.case == "precondition"
    with pytest.raises(gpu.FFTError) as ei:
        gpu.Convolver(np.ones(8, np.float32), 100)
    assert ei.value.case == "unsupportedFFTSize"
    assert gpu.Convolver.recommended_fft_size(1023) == 1024
    assert gpu.Convolver.recommended_fft_size(1024) == 2048
    assert gpu.Convolver.recommended_fft_size(16384) is None


def test_batched_device_apply(gpu, oracle):
    import torch
    count, S, K = 64, 20000, 1023
    x = rnd((count, S), 4)
    h = rnd(K, 5, 1 / np.sqrt(K))
    cv = gpu.Convolver(h, 4096)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty((count, S + K - 1), device="cuda")
    cv.apply_device(xd, yd, S, count)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    for i in (0, 17, 63):
        assert oracle.rel_l2(y[i], oracle.f64_convolve(x[i], h)) < 1e-6
    assert np.array_equal(y[17], cv.apply(x[17]))  # batch == single, bit-exact


class FusedConvolve:
    """torch.autograd.Function over mifft_convolver_{apply,backward}_device."""

    @staticmethod
    def make(gpu):
        import torch

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, x, taps, conv):
                count, S = x.shape
                K = taps.numel()
                conv.set_taps_device(taps)
                y = torch.empty((count, S + K - 1), device=x.device, dtype=torch.float32)
                conv.apply_device(x, y, S, count)
                ctx.save_for_backward(x, taps)
                ctx.conv = conv
                return y

            @staticmethod
            def backward(ctx, gy):
                x, taps = ctx.saved_tensors
                count, S = x.shape
                gx = torch.empty_like(x)
                gh = torch.empty_like(taps)
                ctx.conv.set_taps_device(taps)
                ctx.conv.backward_device(x, gy.contiguous(), gx, gh, S, count)
                return gx, gh, None

        return Fn


def test_adjoint_small_vs_fp64_and_finite_differences(gpu, oracle):
    import torch
    count, S, K, n = 3, 700, 33, 256
    x = rnd((count, S), 10)
    h = rnd(K, 11, 0.2)
    target_h = rnd(K, 12, 0.2)
    target = np.stack([oracle.f64_convolve(r, target_h) for r in x])

    def loss64(xv, hv):
        y = np.stack([oracle.f64_convolve(r, hv) for r in xv])
        return np.mean((y - target) ** 2)

    y64 = np.stack([oracle.f64_convolve(r, h) for r in x])
    gy64 = 2 * (y64 - target) / y64.size
    gx64, gh64 = oracle.f64_conv_grads(x, h, gy64)
    # pin the FP64 gradient formula with central differences (FP64, eps 1e-4)
    for idx in [(0, 0), (1, 350), (2, 699)]:
        xp = x.astype(np.float64).copy(); xp[idx] += 1e-4
        xm = x.astype(np.float64).copy(); xm[idx] -= 1e-4
        fd = (loss64(xp, h) - loss64(xm, h)) / 2e-4
        assert abs(fd - gx64[idx]) < 1e-6 * max(1.0, abs(fd)) + 1e-9
    for k in (0, 16, 32):
        hp = h.astype(np.float64).copy(); hp[k] += 1e-4
        hm = h.astype(np.float64).copy(); hm[k] -= 1e-4
        fd = (loss64(x, hp) - loss64(x, hm)) / 2e-4
        assert abs(fd - gh64[k]) < 1e-6 * max(1.0, abs(fd)) + 1e-9

    Fn = FusedConvolve.make(gpu)
    conv = gpu.Convolver(h, n)
    xd = torch.from_numpy(x).cuda().requires_grad_(True)
    hd = torch.from_numpy(h).cuda().requires_grad_(True)
    td = torch.from_numpy(target.astype(np.float32)).cuda()
    y = Fn.apply(xd, hd, conv)
    loss = torch.mean((y - td) ** 2)
    loss.backward()
    assert oracle.rel_l2(xd.grad.cpu().numpy(), gx64) < 1e-5
    assert oracle.rel_l2(hd.grad.cpu().numpy(), gh64) < 1e-5
    # central differences through the GPU forward itself (FP32, coarse eps)
    with torch.no_grad():
        for k in (3, 20):
            hp = hd.detach().clone(); hp[k] += 1e-2
            hm = hd.detach().clone(); hm[k] -= 1e-2
            lp = torch.mean((Fn.apply(xd.detach(), hp, conv) - td) ** 2).item()
            lm = torch.mean((Fn.apply(xd.detach(), hm, conv) - td) ** 2).item()
            fd = (lp - lm) / 2e-2
            assert abs(fd - hd.grad[k].item()) < 2e-2 * abs(hd.grad[k].item()) + 1e-6


def test_adjoint_identity(gpu):
    """<conv(x, h), g> == <x, A_h^T g> == <h, A_x^T g> in FP32 (sums in FP64)."""
    import torch
    count, S, K = 8, 30000, 1023
    x = rnd((count, S), 20)
    h = rnd(K, 21, 1 / np.sqrt(K))
    g = rnd((count, S + K - 1), 22)
    cv = gpu.Convolver(h, 4096)
    xd, gd = torch.from_numpy(x).cuda(), torch.from_numpy(g).cuda()
    yd = torch.empty_like(gd)
    gx = torch.empty_like(xd)
    gh = torch.empty(K, device="cuda")
    cv.apply_device(xd, yd, S, count)
    cv.backward_device(xd, gd, gx, gh, S, count)
    torch.cuda.synchronize()
    lhs = float(np.sum(yd.cpu().numpy().astype(np.float64) * g))
    rx = float(np.sum(x.astype(np.float64) * gx.cpu().numpy()))
    rh = float(np.sum(h.astype(np.float64) * gh.cpu().numpy()))
    assert abs(lhs - rx) < 1e-5 * abs(lhs)
    assert abs(lhs - rh) < 1e-5 * abs(lhs)


def test_config4_forward_backward(gpu, oracle):
    """512 signals x 65536 ~N(0,1), taps K=1023 ~N(0,1/1023), FFT 4096; loss =
    MSE to the output of a hidden seed-3 filter."""
    import torch
    count, S, K = 512, 65536, 1023
    g0 = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((count, S), generator=g0, device="cuda")
    h = (torch.randn(K, generator=torch.Generator(device="cuda").manual_seed(1), device="cuda")
         / np.sqrt(K)).requires_grad_(True)
    hidden = torch.randn(K, generator=torch.Generator(device="cuda").manual_seed(3),
                         device="cuda") / np.sqrt(K)
    conv = gpu.FusedConvolver(h.detach().cpu().numpy(), 4096)
    target = torch.empty((count, S + K - 1), device="cuda")
    conv.set_taps_device(hidden)
    conv.apply_device(x, target, S, count)
    Fn = FusedConvolve.make(gpu)
    xr = x.clone().requires_grad_(True)
    y = Fn.apply(xr, h, conv)
    loss = torch.mean((y - target) ** 2)
    loss.backward()
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    hs = h.detach().cpu().numpy()
    rows = [0, 255, 511]
    for r in rows:
        assert oracle.rel_l2(y[r].detach().cpu().numpy(), oracle.f64_convolve(xs[r], hs)) < 1e-6
    gy = (2.0 / y.numel()) * (y - target).detach()
    gy64 = gy.cpu().numpy().astype(np.float64)
    gx64, gh64 = oracle.f64_conv_grads(xs, hs, gy64)
    assert oracle.rel_l2(xr.grad.cpu().numpy()[rows], gx64[rows]) < 1e-5
    assert oracle.rel_l2(h.grad.cpu().numpy(), gh64) < 1e-5


def test_correlator(gpu, oracle):
    rng = np.random.default_rng(30)
    for n in (64, 1024, 16384):
        c = gpu.Correlator(n)
        a = rng.standard_normal(n + 37).astype(np.float32)   # truncated (Correlate.swift:54)
        b = rng.standard_normal(n // 3).astype(np.float32)   # zero-padded
        r = c.cross(a, b)
        assert oracle.rel_l2(r, oracle.f64_correlate(a, b, n)) < 1e-6
        assert oracle.rel_l2(r, oracle.cpu_correlate(a, b, n)) < 1e-6
        x = rng.standard_normal(n // 2).astype(np.float32)
        ac = c.auto(x)
        assert abs(ac[0] - np.sum(x.astype(np.float64) ** 2)) < 1e-4 * ac[0]
        assert int(np.argmax(ac)) == 0


@pytest.mark.parametrize("n,count", [(4096, 300), (4096, 301), (4096, 2500), (256, 7), (64, 1)])
def test_correlator_device_batch(gpu, oracle, n, count):
    """Batched device correlations (fewer and more pairs than resident
    workgroups; odd counts leave the last transform slot one pair, and at
    256 and 64 points one workgroup holds several slots), a truncated
    (n + 37 samples) and b zero-padded (n / 3), against the FP64 lag sum;
    the row after the last pair is never written."""
    import torch
    na, nb = n + 37, n // 3
    rng = np.random.default_rng(31)
    a = rng.standard_normal((count, na)).astype(np.float32)
    b = rng.standard_normal((count, nb)).astype(np.float32)
    c = gpu.Correlator(n)
    out = torch.full((count + 1, n), float("nan"), device="cuda")
    c.cross_device(torch.from_numpy(a).cuda(), na, torch.from_numpy(b).cuda(), nb, count, out)
    torch.cuda.synchronize()
    assert torch.isnan(out[count]).all()
    r = out[:count].cpu().numpy()
    for i in sorted({0, 1, count // 2, count - 1} & set(range(count))):
        assert oracle.rel_l2(r[i], oracle.f64_correlate(a[i], b[i], n)) < 1e-6, i
    # every row against the host FFT in FP64 (truncate / zero-pad, IFFT(A conj(B)))
    ap = a[:, :n].astype(np.float64)
    bp = np.zeros((count, n))
    bp[:, :nb] = b
    ref = np.real(np.fft.ifft(np.fft.fft(ap, axis=1) * np.conj(np.fft.fft(bp, axis=1)), axis=1))
    assert oracle.rel_l2(r, ref) < 1e-6
    c.close()


def test_correlator_lag_convention(gpu):
    # r[m] = sum_n a[n+m] b[n]: a delayed copy of b peaks at the delay
    n = 1024
    b = np.zeros(n, np.float32)
    b[100:110] = 1
    a = np.roll(b, 37)
    r = gpu.Correlator(n).cross(a, b)
    assert int(np.argmax(r)) == 37


FP16_MODES = [("float16Pure", "pure", 2e-2), ("float16Storage", "storage", 1e-3),
              ("float16Mixed", "mixed", 2e-2)]


@pytest.mark.parametrize("name,mode,thr", FP16_MODES)
def test_fp16_modes_vs_fp32_and_reference_emulation(gpu, oracle, name, mode, thr):
    # FusedConvolverTests.swift:39-58: 64 taps, fftSize + 500 samples, the
    # FP16 result within `thr` relative L2 of the FP32 fused one; and no
    # worse than twice the error of the reference's own FP16 algorithm
    # (oracle emulation, numpy float16) against FP64
    rng = np.random.default_rng(11)
    taps = rng.uniform(-1, 1, 64).astype(np.float32)
    sig = rng.uniform(-1, 1, 4096 + 500).astype(np.float32)
    fp32 = gpu.FusedConvolver(taps, 4096).apply(sig)
    cv = gpu.FusedConvolver(taps, 4096, getattr(gpu.Precision, name))
    out = cv.apply(sig)
    assert out.size == sig.size + 63
    assert np.all(np.isfinite(out))
    assert oracle.rel_l2(out, fp32) < thr
    ref64 = oracle.f64_convolve(sig, taps)
    emu = oracle.fused_convolve_fp16_emulated(sig, taps, mode)
    err, err_ref = oracle.rel_l2(out, ref64), oracle.rel_l2(emu, ref64)
    assert err < 2 * err_ref, (err, err_ref)
    assert err > 1e-6  # FP16 really was used


@pytest.mark.parametrize("name,mode,thr", FP16_MODES)
def test_fp16_modes_batched_device(gpu, oracle, name, mode, thr):
    import torch
    rng = np.random.default_rng(12)
    taps = (rng.standard_normal(1023) / np.sqrt(1023)).astype(np.float32)
    x = rng.standard_normal((8, 20000)).astype(np.float32)
    cv = gpu.FusedConvolver(taps, 4096, getattr(gpu.Precision, name))
    xs = torch.from_numpy(x).cuda()
    ys = torch.empty((8, 20000 + 1022), dtype=torch.float32, device="cuda")
    cv.apply_device(xs, ys, 20000, 8)
    torch.cuda.synchronize()
    y = ys.cpu().numpy()
    for r in range(8):
        ref64 = oracle.f64_convolve(x[r], taps)
        emu = oracle.fused_convolve_fp16_emulated(x[r], taps, mode)
        assert oracle.rel_l2(y[r], ref64) < 2 * oracle.rel_l2(emu, ref64)


def test_fp16_adjoint_not_built(gpu):
    import torch
    cv = gpu.FusedConvolver(np.ones(8, np.float32), 4096, gpu.Precision.float16Storage)
    x = torch.zeros((1, 100), device="cuda")
    with pytest.raises(gpu.FFTError) as ei:
        cv.backward_device(x, torch.zeros((1, 107), device="cuda"), torch.zeros_like(x),
                           torch.zeros(8, device="cuda"), 100, 1)
    assert ei.value.case == "kernelNotFound"
