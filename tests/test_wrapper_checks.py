"""Host-side argument checks of the Python device wrappers (no GPU).

Every *_device wrapper checks each torch tensor it hands to the C-ABI --
on the GPU, contiguous, of the element type the entry point reads or writes,
and at least as large as the counts passed imply -- and raises FFTError
(invalidArgument) before the library is called.  The objects are built
without their constructors (no device here) and the library functions are
replaced by recorders, so a call that reaches the C-ABI is visible."""
import pytest

import metalfft_amd as m
from metalfft_amd import _lib, api, dsp, sar


class FakeTensor:
    """Enough of torch.Tensor for the wrappers' checks."""

    def __init__(self, numel, dtype="torch.float32", esize=4, cuda=True, contiguous=True):
        self._numel, self.dtype, self._esize = numel, dtype, esize
        self.is_cuda, self._contig = cuda, contiguous

    def numel(self):
        return self._numel

    def element_size(self):
        return self._esize

    def is_contiguous(self):
        return self._contig

    def is_complex(self):
        return "complex" in self.dtype

    def data_ptr(self):
        return 0x1000


def c64(n):
    return FakeTensor(n, "torch.complex64", 8)


def f32(n):
    return FakeTensor(n, "torch.float32", 4)


@pytest.fixture
def calls(monkeypatch):
    """Replaces the loaded library by an object whose every function records
    its call and returns MIFFT_OK."""
    seen = []

    class Rec:
        def __getattr__(self, name):
            def fn(*args):
                seen.append(name)
                return 0
            return fn

    monkeypatch.setattr(_lib, "load", lambda: Rec())
    return seen


def bare(cls, **attrs):
    obj = object.__new__(cls)
    for k, v in attrs.items():
        setattr(obj, k, v)
    return obj


def test_execute_checks_sizes_and_types(calls):
    f = bare(api.MetalFFT, _plan=None, size=1024)
    f.execute(c64(4 * 1024), c64(4 * 1024), batch=4)
    assert calls == ["mifft_execute_device"]
    with pytest.raises(m.FFTError) as ei:
        f.execute(c64(4 * 1024), c64(3 * 1024), batch=4)  # dst too small
    assert ei.value.case == "invalidArgument"
    assert ei.value.expected == 8 * 4 * 1024
    with pytest.raises(m.FFTError):
        f.execute(c64(8 * 1024), c64(8 * 1024), batch=9)  # batch beyond the tensors
    with pytest.raises(m.FFTError):
        f.execute(FakeTensor(8 * 1024, "torch.float64", 8), c64(8 * 1024), batch=1)
    with pytest.raises(m.FFTError):
        f.execute(FakeTensor(1024, "torch.complex64", 8, cuda=False), c64(1024), batch=1)
    with pytest.raises(m.FFTError):
        f.execute(FakeTensor(1024, "torch.complex64", 8, contiguous=False), c64(1024), batch=1)
    with pytest.raises(m.FFTError):
        f.execute(c64(1024), c64(1024), batch=-1)
    # float32 re/im pairs are a complex64 buffer too
    f.execute(f32(2 * 1024), f32(2 * 1024), batch=1)
    assert calls == ["mifft_execute_device"] * 2


def test_real_entry_points(calls):
    f = bare(api.MetalFFT, _plan=None, size=256)
    f.forward_real(f32(2 * 256), c64(2 * 129), batch=2, half=True)
    with pytest.raises(m.FFTError):
        f.forward_real(f32(2 * 256), c64(2 * 129), batch=2)  # all bins need 2*256
    with pytest.raises(m.FFTError):
        f.forward_real(c64(2 * 256), c64(2 * 256), batch=2)  # real input must be float32
    f.inverse_real(c64(3 * 129), f32(3 * 256), batch=3, half=True)
    with pytest.raises(m.FFTError):
        f.inverse_real(c64(3 * 129), f32(3 * 255), batch=3, half=True)
    assert calls == ["mifft_forward_real_device", "mifft_inverse_real_device"]


def test_ragged_uses_extent(calls):
    r = bare(api.RaggedFFT, _p=None, extent=5000)
    r.execute(c64(5000), c64(5000))
    with pytest.raises(m.FFTError):
        r.execute(c64(5000), c64(4999))
    assert calls == ["mifft_ragged_execute_device"]


def test_precision_plan_io_types(calls):
    p = bare(api.PrecisionFFT, _plan=None, size=4096, precision=3, io=1)
    p.execute(FakeTensor(2 * 4096, "torch.float16", 2), FakeTensor(2 * 4096, "torch.float16", 2), 1)
    with pytest.raises(m.FFTError):  # complex-FP16 plan given complex64 data
        p.execute(c64(4096), c64(4096), 1)
    p.io = 0
    p.execute(c64(4096), c64(4096), 1)
    assert calls == ["mifft_execute_device"] * 2


def test_convolver_and_correlator(calls):
    cv = bare(dsp.Convolver, _cv=None, kernelLen=7)
    cv.apply_device(f32(3 * 100), f32(3 * 106), samples=100, count=3)
    with pytest.raises(m.FFTError):
        cv.apply_device(f32(3 * 100), f32(3 * 105), samples=100, count=3)
    with pytest.raises(m.FFTError):
        cv.set_taps_device(f32(6))
    cv.backward_device(f32(200), f32(212), f32(200), f32(7), samples=100, count=2)
    with pytest.raises(m.FFTError):
        cv.backward_device(f32(200), f32(212), f32(200), f32(6), samples=100, count=2)
    c = bare(dsp.Correlator, _c=None, fftSize=64)
    c.cross_device(f32(2 * 50), 50, f32(2 * 60), 60, 2, f32(2 * 64))
    with pytest.raises(m.FFTError):
        c.cross_device(f32(2 * 50), 50, f32(60), 60, 2, f32(2 * 64))
    assert calls == ["mifft_convolver_apply_device", "mifft_convolver_backward_device",
                     "mifft_correlator_cross_device"]


def test_cross_spectral_outputs(calls, monkeypatch):
    x = bare(dsp.CrossSpectral, _x=None, fftSize=64, channels=3, pairs=3)
    monkeypatch.setattr(dsp.CrossSpectral, "frames", lambda self, s, h: 4)
    x.compute_device(f32(3 * 300), 300, 64, None, f32(3 * 64), c64(3 * 64), f32(3 * 64))
    with pytest.raises(m.FFTError):
        x.compute_device(f32(3 * 300), 300, 64, None, f32(3 * 64), c64(2 * 64), f32(3 * 64))
    with pytest.raises(m.FFTError):
        x.compute_device(f32(2 * 300), 300, 64, None, f32(3 * 64), c64(3 * 64), f32(3 * 64))
    x.stft_device(f32(3 * 300), 300, 64, f32(64), c64(3 * 4 * 64))
    with pytest.raises(m.FFTError):
        x.stft_device(f32(3 * 300), 300, 64, f32(64), c64(3 * 4 * 64), power=f32(64))
    with pytest.raises(m.FFTError):
        x.welch_device(f32(3 * 300), 300, 64, None, 1000.0, f32(2 * 64))
    assert calls == ["mifft_xspec_device", "mifft_stft_device"]


def test_range_doppler_frames(calls):
    p = sar.RdaParams()
    p.n_range, p.n_azimuth, p.io = 512, 256, sar.IO_C32H
    r = bare(sar.RangeDoppler, _r=None, params=p)
    half = lambda n: FakeTensor(n, "torch.float16", 2)  # noqa: E731
    r.process(half(2 * 512 * 256), half(2 * 512 * 256))
    with pytest.raises(m.FFTError):
        r.process(half(2 * 512 * 255), half(2 * 512 * 256))
    with pytest.raises(m.FFTError):  # complex64 data for a complex-FP16 processor
        r.corner_turn(c64(512 * 256), c64(512 * 256))
    p.io = sar.IO_C64
    r.azimuth_compress(c64(512 * 256), c64(512 * 256))
    assert calls == ["mifft_rda_process_device", "mifft_rda_azimuth_compress_device"]
