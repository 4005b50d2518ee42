"""Parity of the HIP FFT (through the C-ABI) with the CPU oracle and FP64.

Bars: FP32 results within 1e-6 relative L2 of FP64 and of the FP32 CPU
restatement of the reference algorithm; the reference's own tolerances
(1e-4 rel-L2, 1e-3 impulse) are looser and implied.
"""
import numpy as np
import pytest

from conftest import normal_c64, uniform_c64

pytestmark = pytest.mark.gpu

SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)


@pytest.fixture(scope="module")
def plans(gpu):
    return {n: gpu.MetalFFT(n) for n in SIZES}


@pytest.mark.parametrize("n", SIZES)
def test_every_variant_matches_fp64_and_cpu(gpu, plans, oracle, n):
    import torch
    x = uniform_c64((37, n), 11 + n)  # ragged-looking batch count exercises tail workgroups
    ref = oracle.f64_fft(x)
    cpu = oracle.cpu_fft(x, threads=8)
    fft = plans[n]
    chosen = fft.kernel_name
    try:
        for name in gpu.kernel_variants(n):
            fft.select_kernel(name)
            src = torch.from_numpy(x).cuda()
            dst = torch.empty_like(src)
            fft.execute(src, dst, batch=37)
            torch.cuda.synchronize()
            y = dst.cpu().numpy()
            assert oracle.rel_l2(y, ref) < 1e-6, name
            assert oracle.rel_l2(y, cpu) < 1e-6, name
            # inverse round trip through the same variant
            back = torch.empty_like(src)
            fft.execute(dst, back, batch=37, direction=gpu.INVERSE)
            torch.cuda.synchronize()
            assert oracle.rel_l2(back.cpu().numpy(), x) < 1e-6, name
            assert oracle.rel_l2(back.cpu().numpy(), oracle.cpu_fft(cpu, inverse=True)) < 1e-6
    finally:
        fft.select_kernel(chosen)


@pytest.mark.parametrize("n", SIZES)
def test_impulse_all_sizes(plans, n):
    # Tests/MetalFFTTests/MetalFFTTests.swift:11-23
    x = np.zeros(n, np.complex64)
    x[0] = 1
    y = plans[n].forward(x)
    assert np.max(np.abs(y.real - 1)) < 1e-6 and np.max(np.abs(y.imag)) < 1e-6


def test_sinusoid_peak_bin(plans):
    # MetalFFTTests.swift:37-49
    n = 1024
    k = np.arange(n)
    x = np.cos(2 * np.pi * 100 * k / n).astype(np.complex64)
    y = plans[n].forward(x)
    assert int(np.argmax(np.abs(y[: n // 2]))) == 100


def test_golden_vectors(plans, golden, oracle):
    for n in SIZES:
        y = plans[n].forward(golden[f"x_{n}"])
        assert oracle.rel_l2(y, golden[f"X_{n}"]) < 1e-6


def test_batch_matches_single(plans):
    # MetalFFTTests.swift:51-65, here at every size and bit-exact
    for n in SIZES:
        a, b = uniform_c64((2, n), n + 1)
        batch = plans[n].forward_batch([a, b])
        assert np.array_equal(batch[0], plans[n].forward(a))
        assert np.array_equal(batch[1], plans[n].forward(b))


def test_inverse_host_round_trip(plans, oracle):
    # CrossSpectralTests.swift:49-63
    for n in SIZES:
        x = uniform_c64((n,), 3 * n)
        assert oracle.rel_l2(plans[n].inverse(plans[n].forward(x)), x) < 1e-6


def test_forward_into_reuses_output(plans):
    n = 512
    x = uniform_c64((n,), 5)
    out = np.zeros(n, np.complex64)
    got = plans[n].forward_into(x, out)
    assert got is out and np.array_equal(out, plans[n].forward(x))
    resized = plans[n].forward_into(x, np.zeros(3, np.complex64))
    assert resized.size == n


def test_plain_forward_workload_n4096_b256(gpu, plans, oracle, golden):
    """N=4096, batch 256, U[-1,1) seed 0: vs FP64 and the CPU path, then
    inverse round trip."""
    import torch
    x = uniform_c64((256, 4096), 0)
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    plans[4096].execute(src, dst)
    back = torch.empty_like(src)
    plans[4096].execute(dst, back, direction=gpu.INVERSE)
    torch.cuda.synchronize()
    y = dst.cpu().numpy()
    assert oracle.rel_l2(y, oracle.f64_fft(x)) < 1e-6
    assert oracle.rel_l2(y, oracle.cpu_fft(x, threads=8)) < 1e-6
    assert oracle.rel_l2(y[0], golden["cfg1_first_row"]) < 1e-6
    assert np.allclose(np.sum(np.abs(y.astype(np.complex128)) ** 2, -1), golden["cfg1_energy"],
                       rtol=1e-5)
    assert oracle.rel_l2(back.cpu().numpy(), x) < 1e-6


def test_metric_size_n4096_b4096(gpu, plans, oracle):
    """Full metric size: FP64 check on a strided sample of transforms, plus
    size-independent properties on all 4096 (Parseval, linearity, round trip)."""
    import torch
    n, b = 4096, 4096
    x = uniform_c64((b, n), 0)
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    fft = plans[n]
    fft.execute(src, dst)
    torch.cuda.synchronize()
    y = dst.cpu().numpy()
    idx = np.arange(0, b, 61)
    assert oracle.rel_l2(y[idx], oracle.f64_fft(x[idx])) < 1e-6
    # Parseval per transform: sum|X|^2 = N sum|x|^2
    ex = np.sum(np.abs(x.astype(np.complex128)) ** 2, -1) * n
    ey = np.sum(np.abs(y.astype(np.complex128)) ** 2, -1)
    assert np.max(np.abs(ey / ex - 1)) < 1e-5
    # linearity: F(2x + x_shift) == 2 F(x) + F(x_shift)
    x2 = torch.roll(src, 1, 0)
    comb = 2 * src + x2
    fc = torch.empty_like(comb)
    fft.execute(comb, fc)
    f2 = torch.empty_like(x2)
    fft.execute(x2, f2)
    torch.cuda.synchronize()
    lin = (2 * dst + f2).cpu().numpy()
    assert oracle.rel_l2(fc.cpu().numpy(), lin) < 1e-6
    # in-place inverse round trip over the whole batch
    fft.execute(dst, dst, direction=gpu.INVERSE)
    torch.cuda.synchronize()
    assert oracle.rel_l2(dst.cpu().numpy(), x) < 1e-6


def test_in_place_matches_out_of_place(gpu, plans):
    import torch
    for n in (64, 4096, 16384):
        x = torch.from_numpy(normal_c64((9, n), n)).cuda()
        out = torch.empty_like(x)
        plans[n].execute(x, out)
        y = x.clone()
        plans[n].execute(y, y)
        torch.cuda.synchronize()
        assert torch.equal(out, y)


@pytest.mark.parametrize("n", [8192, 16384])
def test_in_place_every_variant(gpu, plans, n):
    """Every kernel form of the long lengths (one-shot, persistent with a
    prefetched next transform, DMA-staged, split-LDS exchange) gives the
    same bits in place as out of place, forward and inverse."""
    import torch
    fft = plans[n]
    chosen = fft.kernel_name
    x = torch.from_numpy(normal_c64((11, n), 3 * n)).cuda()
    try:
        for name in gpu.kernel_variants(n):
            fft.select_kernel(name)
            for direction in (gpu.FORWARD, gpu.INVERSE):
                out = torch.empty_like(x)
                fft.execute(x, out, batch=11, direction=direction)
                y = x.clone()
                fft.execute(y, y, batch=11, direction=direction)
                torch.cuda.synchronize()
                assert torch.equal(out, y), (name, direction)
    finally:
        fft.select_kernel(chosen)


def test_error_cases(gpu, plans):
    # MetalFFT.swift:70-72 and :110-114
    with pytest.raises(gpu.FFTError) as ei:
        plans[256].forward(np.zeros(255, np.complex64))
    assert ei.value.case == "invalidInputSize"
    assert (ei.value.expected, ei.value.got) == (256, 255)
    good = np.zeros(256, np.complex64)
    with pytest.raises(gpu.FFTError) as ei:
        plans[256].forward_batch([good, good, np.zeros(128, np.complex64)])
    assert ei.value.case == "batchInputSize" and ei.value.batch_index == 2
    assert plans[256].forward_batch([]) == []
    with pytest.raises(gpu.FFTError) as ei:
        plans[256].select_kernel("stockham_r16_4096")
    assert ei.value.case == "kernelNotFound"


def test_multichannel(gpu, oracle):
    # Tests/MetalFFTTests/CrossSpectralTests.swift:11-63
    mc = gpu.MultiChannelFFT(4, 4096)
    xs = [uniform_c64((4096,), 40 + c) for c in range(4)]
    ys = mc.forward(xs)
    for x, y in zip(xs, ys):
        assert oracle.rel_l2(y, oracle.f64_fft(x)) < 1e-6
    back = mc.inverse(ys)
    for x, z in zip(xs, back):
        assert oracle.rel_l2(z, x) < 1e-6
    with pytest.raises(gpu.FFTError) as ei:
        mc.forward(xs[:3])
    assert ei.value.case == "precondition"


def test_twiddle_generation_matches_fp64(gpu, oracle):
    """The device's FP64 sincospi table equals the CPU's FP64 table rounded
    to FP32 (impulse at 1 returns W_N^k exactly up to table rounding)."""
    for n in (4096, 16384):
        x = np.zeros(n, np.complex64)
        x[1] = 1
        y = gpu.MetalFFT(n).forward(x)
        tw = oracle.twiddle_table(n)
        assert np.max(np.abs(y - tw)) < 4e-7


def test_diagnostic_copy(gpu):
    """mifft_copy_device (the bench's HBM copy ceiling) copies exactly, full
    blocks, a partial last block and a byte tail, and rejects bad sizes."""
    import ctypes
    import torch
    from metalfft_amd import _lib
    L = _lib.load()
    for nbytes in (16, 32768 * 3, 32768 * 3 + 4096 + 48, 1 << 24, (1 << 20) + 7):
        src = torch.randint(0, 255, (nbytes,), dtype=torch.uint8, device="cuda")
        dst = torch.zeros_like(src)
        rc = L.mifft_copy_device(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                 ctypes.c_int64(nbytes), None)
        assert rc == 0
        torch.cuda.synchronize()
        assert torch.equal(src, dst), nbytes
    assert L.mifft_copy_device(None, None, ctypes.c_int64(-1), None) != 0
    assert L.mifft_copy_device(None, None, ctypes.c_int64(0), None) == 0
