"""The adjoint that allocates nothing (mifft_convolver_backward_device_noalloc,
through the C-ABI) against the allocating adjoint
(mifft_convolver_backward_device) on the same inputs: two calls in a row,
the second smaller, bit for bit where the allocating form's partials fit the
fixed scratch; where they do not (K = 16383 at fft size 16384), grad_signal
bit for bit and grad_taps within FP32 rounding of the other summation order."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_both(gpu, cv, x, gy, samples, count):
    import torch
    from metalfft_amd import _lib
    L = _lib.load()
    K = cv.kernelLen
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = []
    for fn in (L.mifft_convolver_backward_device, L.mifft_convolver_backward_device_noalloc):
        gx = torch.full((count, samples), float("nan"), device="cuda")
        gh = torch.full((K,), float("nan"), device="cuda")
        _lib.check(fn(cv._cv, ctypes.c_void_p(x.data_ptr()), samples, count,
                      ctypes.c_void_p(gy.data_ptr()), ctypes.c_void_p(gx.data_ptr()),
                      ctypes.c_void_p(gh.data_ptr()), st))
        torch.cuda.synchronize()
        outs.append((gx, gh))
    return outs


def test_noalloc_matches_allocating_adjoint(gpu):
    import torch
    rng = np.random.default_rng(31)
    K, N = 1023, 4096
    cv = gpu.FusedConvolver((rng.standard_normal(K) / np.sqrt(K)).astype(np.float32), N)
    for count, samples in ((64, 65536), (8, 20000)):  # second call smaller
        print(f"noalloc adjoint count={count} samples={samples} K={K} N={N}", flush=True)
        x = torch.randn((count, samples), device="cuda")
        gy = torch.randn((count, samples + K - 1), device="cuda")
        (gx0, gh0), (gx1, gh1) = run_both(gpu, cv, x, gy, samples, count)
        assert torch.equal(gx1, gx0) and torch.equal(gh1, gh0)


def test_noalloc_smaller_split(gpu):
    """Partials of the allocating split exceed the fixed scratch: fewer
    workgroups, same grad_signal bits, grad_taps to FP32 rounding."""
    import torch
    rng = np.random.default_rng(32)
    N = 16384
    K = N - 1
    cv = gpu.Convolver((rng.standard_normal(K) / np.sqrt(K)).astype(np.float32), N)
    for count, samples in ((2, 20000), (1, 9000)):
        print(f"noalloc adjoint count={count} samples={samples} K={K} N={N}", flush=True)
        x = torch.randn((count, samples), device="cuda")
        gy = torch.randn((count, samples + K - 1), device="cuda")
        (gx0, gh0), (gx1, gh1) = run_both(gpu, cv, x, gy, samples, count)
        assert torch.equal(gx1, gx0)
        err = (torch.linalg.norm(gh1 - gh0) / torch.linalg.norm(gh0)).item()
        assert err < 1e-6, err


def test_noalloc_zero_sizes_and_errors(gpu):
    import torch
    from metalfft_amd import _lib
    K = 31
    cv = gpu.Convolver(np.ones(K, np.float32), 256)
    gh = torch.full((K,), float("nan"), device="cuda")
    L = _lib.load()
    # no samples: grad_taps is all zeros (the other buffers are not read)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.mifft_convolver_backward_device_noalloc(cv._cv, None, 0, 4, None, None,
                                                         ctypes.c_void_p(gh.data_ptr()), st))
    torch.cuda.synchronize()
    assert not gh.any()
    assert L.mifft_convolver_backward_device_noalloc(None, None, 1, 1, None, None, None, None) == 11
    h16 = gpu.FusedConvolver(np.ones(K, np.float32), precision=gpu.Precision.float16Pure)
    assert L.mifft_convolver_backward_device_noalloc(h16._cv, None, 1, 1, None, None,
                                                     ctypes.c_void_p(gh.data_ptr()), None) == 4
