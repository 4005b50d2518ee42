"""Real-input transforms (mifft_forward_real_device / mifft_inverse_real_device)
on the GPU through the C-ABI: the spectra the reference's callers get from
forward() on (x, 0)-packed real signals (Convolve.swift:48-53,
FusedConvolver.swift:101-107, Correlate.swift:53-61), against FP64 and the
oracle's FP32 restatement of the complex path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)


def rnd(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape).astype(np.float32)


@pytest.mark.parametrize("n", SIZES)
def test_forward_real_vs_fp64_and_restatement(gpu, oracle, n):
    import torch
    batch = 37
    x = rnd((batch, n), n)
    f = gpu.MetalFFT(n)
    dx = torch.from_numpy(x).cuda()
    full = torch.empty((batch, n), dtype=torch.complex64, device="cuda")
    half = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda")
    f.forward_real(dx, full, batch)
    f.forward_real(dx, half, batch, half=True)
    torch.cuda.synchronize()
    want = np.fft.fft(x.astype(np.float64), axis=1)
    got = full.cpu().numpy()
    assert oracle.rel_l2(got, want) < 1e-6
    assert oracle.rel_l2(half.cpu().numpy(), want[:, :n // 2 + 1]) < 1e-6
    assert np.array_equal(half.cpu().numpy(), got[:, :n // 2 + 1])
    packed = oracle.cpu_fft(x.astype(np.complex64))  # the callers' (x, 0) packing
    assert oracle.rel_l2(got, packed) < 2e-6


@pytest.mark.parametrize("n", SIZES)
def test_inverse_real_round_trip(gpu, oracle, n):
    import torch
    batch = 9
    x = rnd((batch, n), n + 1)
    f = gpu.MetalFFT(n)
    X = np.fft.fft(x.astype(np.float64), axis=1)
    for half in (False, True):
        src = X[:, :n // 2 + 1] if half else X
        dy = torch.from_numpy(np.ascontiguousarray(src.astype(np.complex64))).cuda()
        out = torch.empty((batch, n), dtype=torch.float32, device="cuda")
        f.inverse_real(dy, out, batch, half=half)
        torch.cuda.synchronize()
        assert oracle.rel_l2(out.cpu().numpy(), x) < 1e-6


def test_real_zero_and_impulse_exact(gpu):
    """A zero signal gives exact zeros (no transform is shared between
    signals); an impulse at 0 gives all ones; a constant c gives N c at DC."""
    import torch
    n = 1024
    x = np.zeros((3, n), np.float32)
    x[1, 0] = 1.0
    x[2, :] = 0.5
    f = gpu.MetalFFT(n)
    y = torch.empty((3, n), dtype=torch.complex64, device="cuda")
    f.forward_real(torch.from_numpy(x).cuda(), y, 3)
    torch.cuda.synchronize()
    y = y.cpu().numpy()
    assert not y[0].any()
    assert np.max(np.abs(y[1] - 1)) < 1e-6
    assert abs(y[2, 0] - 0.5 * n) < 1e-3 and np.max(np.abs(y[2, 1:])) < 1e-3


def test_forward_real_host_mirror(gpu, oracle):
    n = 2048
    x = rnd(n, 5)
    f = gpu.MetalFFT(n)
    assert oracle.rel_l2(f.forwardReal(x), np.fft.fft(x.astype(np.float64))) < 1e-6
    with pytest.raises(gpu.FFTError) as e:
        f.forwardReal(x[:-1])
    assert e.value.case == "invalidInputSize"


def test_real_large_batch_parseval(gpu):
    """4096 x 4096 real signals: Parseval per signal (sum |X|^2 = N sum x^2)
    and a round trip."""
    import torch
    n, batch = 4096, 4096
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn((batch, n), generator=g, device="cuda")
    f = gpu.MetalFFT(n)
    y = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda")
    f.forward_real(x, y, batch, half=True)
    back = torch.empty_like(x)
    f.inverse_real(y, back, batch, half=True)
    torch.cuda.synchronize()
    p = (y.abs() ** 2).double()
    e = 2 * p.sum(1) - p[:, 0] - p[:, -1]  # bins 1..N/2-1 appear twice
    ref = n * (x.double() ** 2).sum(1)
    assert float(((e - ref).abs() / ref).max()) < 1e-5
    assert float((back - x).norm() / x.norm()) < 1e-6


def test_real_errors(gpu):
    import torch
    f = gpu.MetalFFT(256)
    d = torch.zeros(512, device="cuda")
    from metalfft_amd import _lib
    st = _lib.load().mifft_forward_real_device(f._plan, d.data_ptr(), d.data_ptr(), 1, 2, None)
    assert _lib.STATUS_CASES[st] == "invalidArgument"
    with pytest.raises(gpu.FFTError):
        f.forward_real(d, d, -1)
    f.forward_real(d, d, 0)  # no-op


@pytest.mark.parametrize("n", [64, 2048])
def test_real_unaligned_buffers(gpu, oracle, n):
    """Input and output one float past an 8-byte boundary: the two-load /
    two-store path gives the same spectra and signals."""
    import torch
    batch = 5
    x = rnd((batch, n), 77)
    f = gpu.MetalFFT(n)
    raw = torch.zeros(batch * n + 1, device="cuda")
    raw[1:] = torch.from_numpy(x.reshape(-1)).cuda()
    y = torch.empty((batch, n), dtype=torch.complex64, device="cuda")
    f.forward_real(raw[1:], y, batch)
    back = torch.zeros(batch * n + 1, device="cuda")
    f.inverse_real(y, back[1:], batch)
    torch.cuda.synchronize()
    assert oracle.rel_l2(y.cpu().numpy(), np.fft.fft(x.astype(np.float64), axis=1)) < 1e-6
    assert oracle.rel_l2(back[1:].cpu().numpy().reshape(batch, n), x) < 1e-6
