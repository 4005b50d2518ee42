"""DFT-as-GEMM (MFMA) N=4096 FFT in FP16 / BF16 with FP32 accumulation,
plain and hi/lo split, complex64 or complex-FP16 device data; errors vs FP64
on the sum-of-tones workload (60 dB amplitude range plus noise)."""
import ctypes

import numpy as np
import pytest

from conftest import tones_c64

pytestmark = pytest.mark.gpu

# (precision code, name, rel-L2 bound with complex64 I/O)
PRECISIONS = [(1, "fp16", 2e-3), (2, "bf16", 2e-2), (3, "fp16_split", 1e-5), (4, "bf16_split", 2e-4)]
PRECISIONS_BOUND = [(p, b) for p, _, b in PRECISIONS]


def make_plan(gpu, prec, io):
    from metalfft_amd import _lib
    L = _lib.load()
    p = ctypes.c_void_p()
    _lib.check(L.mifft_plan_create_ex(4096, prec, io, ctypes.byref(p)))
    return p


def run(gpu, plan, src, dst, batch, direction=-1):
    import torch
    from metalfft_amd import _lib
    _lib.check(_lib.load().mifft_execute_device(plan, src.data_ptr(), dst.data_ptr(), batch, direction,
                                                torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def tones():
    return tones_c64(48, 4096, 3)


@pytest.mark.parametrize("prec,name,bound", PRECISIONS)
def test_mfma_c64_vs_fp64(gpu, oracle, tones, prec, name, bound):
    import torch
    plan = make_plan(gpu, prec, 0)
    x = torch.from_numpy(tones).cuda()
    y = torch.empty_like(x)
    run(gpu, plan, x, y, tones.shape[0])
    ref = oracle.f64_fft(tones)
    err = oracle.rel_l2(y.cpu().numpy(), ref)
    print(f"{name} c64 rel_l2={err:.3e} max_abs={oracle.max_abs(y.cpu().numpy(), ref):.3e}")
    assert err < bound
    back = torch.empty_like(x)
    run(gpu, plan, y, back, tones.shape[0], direction=1)
    assert oracle.rel_l2(back.cpu().numpy(), tones) < 2 * bound


@pytest.mark.parametrize("prec,name,bound", PRECISIONS)
def test_mfma_complex_fp16_io(gpu, oracle, tones, prec, name, bound):
    """complex-FP16 in and out: input and output rounding add ~2e-4."""
    import torch
    plan = make_plan(gpu, prec, 1)
    xh = torch.from_numpy(tones.view(np.float32)).cuda().to(torch.float16)  # (b, 2n) re/im
    yh = torch.empty_like(xh)
    run(gpu, plan, xh, yh, tones.shape[0])
    y = yh.float().cpu().numpy().view(np.complex64)
    err = oracle.rel_l2(y, oracle.f64_fft(tones))
    print(f"{name} c32h rel_l2={err:.3e}")
    assert err < max(bound, 1e-3)


def test_mfma_random_and_impulse(gpu, oracle):
    import torch
    from conftest import uniform_c64
    plan = make_plan(gpu, 3, 0)
    x = uniform_c64((16, 4096), 9)
    x[0] = 0
    x[0, 0] = 1  # impulse -> ones
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    run(gpu, plan, xd, yd, 16)
    y = yd.cpu().numpy()
    assert np.max(np.abs(y[0] - 1)) < 1e-5
    assert oracle.rel_l2(y[1:], oracle.f64_fft(x[1:])) < 1e-5
    # zero transform stays zero (scale exponent of an all-zero input)
    xd.zero_()
    run(gpu, plan, xd, yd, 16)
    assert not yd.any()


def test_mfma_needs_n4096(gpu):
    from metalfft_amd import _lib
    p = ctypes.c_void_p()
    assert _lib.load().mifft_plan_create_ex(1024, 3, 0, ctypes.byref(p)) == 4
    assert _lib.load().mifft_plan_create_ex(4096, 0, 1, ctypes.byref(p)) == 4


@pytest.mark.parametrize("prec", [1, 2, 3, 4])
def test_mfma_subnormal_row(gpu, oracle, prec):
    """A transform whose largest value is subnormal (~1e-39) keeps a finite
    power-of-two scale: no inf or NaN, and the rows around it are unaffected."""
    import torch
    from conftest import uniform_c64
    plan = make_plan(gpu, prec, 0)
    x = uniform_c64((4, 4096), 21)
    x[1] = (x[1] * np.float32(1e-39)).astype(np.complex64)
    x[1, 5:] = 0
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    run(gpu, plan, xd, yd, 4)
    y = yd.cpu().numpy()
    assert np.isfinite(y.view(np.float32)).all()
    ref = oracle.f64_fft(x.astype(np.complex128))
    bound = dict(PRECISIONS_BOUND)[prec]
    assert oracle.rel_l2(y[[0, 2, 3]], ref[[0, 2, 3]]) < bound
    # the subnormal row: FP16/BF16 keep ~2^-125 * 2^-24 absolute at worst
    assert np.max(np.abs(y[1] - ref[1])) < 1e-38


# persistent DMA-staged or one transform per workgroup, three radix-16 or
# two radix-64 passes
FORMS = ("", "_oneshot", "_r64", "_r64_oneshot")


@pytest.mark.parametrize("prec,name,bound", PRECISIONS)
@pytest.mark.parametrize("io", [0, 1])
def test_mfma_every_form(gpu, oracle, tones, prec, name, bound, io):
    """Every kernel form of every precision and element type (the autotuner
    picks one per plan; mifft_plan_select_kernel pins any) against FP64,
    and the forms agree with the first to their rounding."""
    import torch
    from metalfft_amd import _lib
    L = _lib.load()
    plan = make_plan(gpu, prec, io)
    chosen = L.mifft_plan_kernel_name(plan).decode()
    base = chosen.replace("_oneshot", "").replace("_r64", "")
    assert base.startswith("mfma_")
    if io == 0:
        x = torch.from_numpy(tones).cuda()
    else:
        x = torch.from_numpy(tones.view(np.float32)).cuda().to(torch.float16)
    ref = oracle.f64_fft(tones)
    outs = []
    for form in FORMS:
        _lib.check(L.mifft_plan_select_kernel(plan, (base + form).encode()))
        y = torch.empty_like(x)
        run(gpu, plan, x, y, tones.shape[0])
        yc = y.cpu().numpy() if io == 0 else y.float().cpu().numpy().view(np.complex64)
        assert oracle.rel_l2(yc, ref) < max(bound, 1e-3 if io else 0), (name, io, form)
        outs.append(yc)
    for o in outs[1:]:
        assert oracle.rel_l2(o, outs[0]) < max(bound, 1e-3 if io else 0)


@pytest.mark.parametrize("prec,name,bound", PRECISIONS)
def test_mfma_full_batch_4096(gpu, oracle, prec, name, bound):
    """The metric's shape, N = 4096 x batch 4096 of the 20-tone workload, in
    the tuned form: FP64 on 32 sampled transforms."""
    import torch
    from conftest import tones_c64
    plan = make_plan(gpu, prec, 0)
    rows = np.arange(0, 4096, 128)
    x = torch.empty((4096, 4096), dtype=torch.complex64, device="cuda")
    sample = tones_c64(len(rows), 4096, 77)
    x.copy_(torch.repeat_interleave(torch.from_numpy(sample).cuda(), 4096 // len(rows), dim=0))
    y = torch.empty_like(x)
    run(gpu, plan, x, y, 4096)
    got = y[torch.from_numpy(rows).cuda()].cpu().numpy()
    src = x[torch.from_numpy(rows).cuda()].cpu().numpy()
    err = oracle.rel_l2(got, oracle.f64_fft(src))
    print(f"{name} batch 4096 rel_l2={err:.3e}")
    assert err < bound
