"""Cross-check against an independent GPU FFT on the same device: torch.fft
(rocFFT through hipFFT on ROCm).  Both are FP32; their difference is within
the sum of their errors against FP64 (each ~1e-7)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)


@pytest.mark.parametrize("n", SIZES)
def test_forward_inverse_vs_rocfft(gpu, n):
    import torch
    batch = max(1, (1 << 20) // n)
    g = torch.Generator(device="cuda").manual_seed(n)
    x = torch.randn((batch, n), dtype=torch.complex64, device="cuda", generator=g)
    f = gpu.MetalFFT(n)
    y = torch.empty_like(x)
    f.execute(x, y, batch)
    z = torch.empty_like(x)
    f.execute(x, z, batch, direction=gpu.INVERSE)
    ref_f = torch.fft.fft(x, dim=1)
    ref_i = torch.fft.ifft(x, dim=1)
    torch.cuda.synchronize()
    assert float((y - ref_f).norm() / ref_f.norm()) < 1e-6
    assert float((z - ref_i).norm() / ref_i.norm()) < 1e-6


@pytest.mark.parametrize("n", (64, 1024, 16384))
def test_real_forward_vs_rocfft(gpu, n):
    import torch
    batch = max(1, (1 << 20) // n)
    g = torch.Generator(device="cuda").manual_seed(n + 1)
    x = torch.randn((batch, n), device="cuda", generator=g)
    f = gpu.MetalFFT(n)
    y = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda")
    f.forward_real(x, y, batch, half=True)
    ref = torch.fft.rfft(x, dim=1)
    back = torch.empty_like(x)
    f.inverse_real(y, back, batch, half=True)
    torch.cuda.synchronize()
    assert float((y - ref).norm() / ref.norm()) < 1e-6
    assert float((back - torch.fft.irfft(ref, n=n, dim=1)).norm() / x.norm()) < 1e-6
