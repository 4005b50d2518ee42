"""Four-step transforms above 16384 points (32768 ... 2^20) on the GPU,
through the C-ABI (at 32768 also the single-workgroup half-LDS kernel, the
autotuner's other candidate): every compiled variant against FP64 and the oracle's
FP32 four-step restatement of MetalFFT.swift:206-277, within 1e-6 rel-L2;
inverse round trip; in-place == out-of-place and batch == single bit for
bit; the host entry points; hipGraph replay."""
import numpy as np
import pytest

from conftest import uniform_c64

pytestmark = pytest.mark.gpu

FOURSTEP_SIZES = (32768, 65536, 131072, 262144, 524288, 1 << 20)


def batch_for(n):
    return 3 if n <= 262144 else 2


@pytest.mark.parametrize("n", FOURSTEP_SIZES)
def test_every_fourstep_variant(gpu, oracle, n):
    import torch
    b = batch_for(n)
    x = uniform_c64((b, n), n + 17)
    ref = oracle.f64_fft(x)
    cpu = oracle.cpu_fft(x, threads=8)
    fft = gpu.MetalFFT(n)
    names = gpu.kernel_variants(n)
    assert any(s.startswith("fourstep_") for s in names)
    assert fft.kernel_name in names
    src = torch.from_numpy(x).cuda()
    for name in names:
        fft.select_kernel(name)
        print(f"fourstep n={n} variant={name}", flush=True)
        dst = torch.empty_like(src)
        fft.execute(src, dst, batch=b)
        torch.cuda.synchronize()
        y = dst.cpu().numpy()
        assert oracle.rel_l2(y, ref) < 1e-6, name
        assert oracle.rel_l2(y, cpu) < 1e-6, name
        back = torch.empty_like(src)
        fft.execute(dst, back, batch=b, direction=gpu.INVERSE)
        torch.cuda.synchronize()
        assert oracle.rel_l2(back.cpu().numpy(), x) < 1e-6, name
        assert oracle.rel_l2(back.cpu().numpy(), oracle.cpu_fft(cpu, inverse=True, threads=8)) < 1e-6


@pytest.mark.parametrize("n", FOURSTEP_SIZES[1:])
def test_cache_policies_same_bits(gpu, n):
    """A shape compiled with another cache policy for its intermediate and
    output (suffix _y<p>o<q>) runs the same arithmetic: forward and inverse
    equal the streaming-policy shape bit for bit, also in place."""
    import re
    import torch
    x = torch.from_numpy(uniform_c64((2, n), n + 5)).cuda()
    fft = gpu.MetalFFT(n)
    names = [s for s in gpu.kernel_variants(n) if s.startswith("fourstep_")]
    policy = [s for s in names if re.search(r"_y\d+o\d+$", s)]
    assert policy
    for name in policy:
        base = re.sub(r"_y\d+o\d+$", "", name)
        assert base in names, name
        outs = []
        for k in (base, name):
            fft.select_kernel(k)
            print(f"fourstep n={n} variant={k}", flush=True)
            y = torch.empty_like(x)
            fft.execute(x, y, batch=2)
            z = x.clone()
            fft.execute(z, z, batch=2, direction=gpu.INVERSE)
            torch.cuda.synchronize()
            outs.append((y, z))
        assert torch.equal(outs[0][0], outs[1][0]), name
        assert torch.equal(outs[0][1], outs[1][1]), name


@pytest.mark.parametrize("n", [32768, 1 << 20])
def test_in_place_and_batch_single(gpu, n):
    import torch
    b = 3
    x = torch.from_numpy(uniform_c64((b, n), 7)).cuda()
    fft = gpu.MetalFFT(n)
    out = torch.empty_like(x)
    fft.execute(x, out, batch=b)
    inplace = x.clone()
    fft.execute(inplace, inplace, batch=b)  # through the plan's scratch
    one = torch.empty_like(x[1])
    fft.execute(x[1], one, batch=1)
    torch.cuda.synchronize()
    assert torch.equal(inplace, out)
    assert torch.equal(one, out[1])
    # inverse in place returns the input
    fft.execute(inplace, inplace, batch=b, direction=gpu.INVERSE)
    torch.cuda.synchronize()
    err = (torch.linalg.norm(inplace - x) / torch.linalg.norm(x)).item()
    assert err < 1e-6


def test_known_answers(gpu, oracle):
    for n in (32768, 262144):
        fft = gpu.MetalFFT(n)
        imp = np.zeros(n, np.complex64)
        imp[0] = 1
        assert np.max(np.abs(fft.forward(imp) - 1)) < 1e-6
        # impulse at p -> W_N^{pk}: pins the transpose order and the twiddle
        p = 3 * n // 7
        imp[0], imp[p] = 0, 1
        k = np.arange(n)
        expect = np.exp(-2j * np.pi * ((p * k) % n) / n)
        assert oracle.max_abs(fft.forward(imp), expect) < 1e-5
        dc = fft.forward(np.ones(n, np.complex64))
        assert abs(dc[0] - n) < 1e-6 * n and np.max(np.abs(dc[1:])) < 1e-6 * n


def test_host_entry_points(gpu, oracle):
    n = 65536
    fft = gpu.MetalFFT(n)
    a, b = uniform_c64((2, n), 99)
    ya = fft.forward(a)
    assert oracle.rel_l2(ya, oracle.f64_fft(a)) < 1e-6
    both = fft.forward_batch([a, b])
    assert np.array_equal(both[0], ya)
    assert oracle.rel_l2(fft.inverse(ya), a) < 1e-6
    with pytest.raises(gpu.FFTError) as ei:
        fft.forward(a[:-1])
    assert ei.value.case == "invalidInputSize"


def test_objects_stay_single_pass(gpu):
    """The convolver, correlator and analysers take FFT sizes up to 16384,
    and real-input entry points do not run on four-step plans."""
    import torch
    with pytest.raises(gpu.FFTError) as ei:
        gpu.Correlator(32768)
    assert ei.value.case == "unsupportedFFTSize"
    with pytest.raises(gpu.FFTError) as ei:
        gpu.Convolver(np.ones(3, np.float32), 32768)
    assert ei.value.case == "unsupportedFFTSize"
    fft = gpu.MetalFFT(32768)
    x = torch.zeros(32768, device="cuda")
    y = torch.zeros(32768, dtype=torch.complex64, device="cuda")
    with pytest.raises(gpu.FFTError) as ei:
        fft.forward_real(x, y, 1)
    assert ei.value.case == "kernelNotFound"


def test_graph_replay(gpu):
    import torch
    n, b = 131072, 4
    x = torch.randn((b, n), dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    z = x.clone()
    fft = gpu.MetalFFT(n)

    def run():
        fft.execute(x, y, b)
        fft.execute(z, z, b)  # in place: the scratch grows in the eager call
    run()
    torch.cuda.synchronize()
    ref_y = y.clone()
    z.copy_(x)
    run()
    torch.cuda.synchronize()
    ref_z = z.clone()
    g = torch.cuda.CUDAGraph()
    z.copy_(x)
    with torch.cuda.graph(g):
        run()
    for r in range(3):
        y.zero_()
        z.copy_(x)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _ = (torch.randn(4096, device="cuda") * r).sum().item()
        assert torch.equal(y, ref_y) and torch.equal(z, ref_z)
    del g
