import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def uniform_c64(shape, seed):
    """re, im i.i.d. U[-1, 1) from numpy default_rng(seed), complex64."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=tuple(shape) + (2,))
    return x.astype(np.float32).view(np.complex64)[..., 0]


def normal_c64(shape, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(size=tuple(shape) + (2,)) * scale
    return x.astype(np.float32).view(np.complex64)[..., 0]


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(ROOT, "tests", "golden", "fft_golden.npz")
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="session")
def gpu():
    """The library on a real GPU; fails (not skips) when the build is missing."""
    import metalfft_amd as m
    from metalfft_amd import _lib
    _lib.load()
    m.device_name()
    return m


def tones_c64(batch, n, seed, tones=20, noise=1e-3):
    """Sum of `tones` complex tones per transform (random frequency and phase,
    amplitudes log-uniform over 60 dB) plus complex N(0, noise^2) noise."""
    rng = np.random.default_rng(seed)
    k = np.arange(n)
    out = np.empty((batch, n), np.complex128)
    for b in range(batch):
        f = rng.uniform(0, n, tones)
        ph = rng.uniform(0, 2 * np.pi, tones)
        amp = 10.0 ** (-3.0 * rng.uniform(0, 1, tones))
        out[b] = (amp[:, None] * np.exp(1j * (2 * np.pi * f[:, None] * k[None, :] / n + ph[:, None]))).sum(0)
    out += noise * (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n)))
    return out.astype(np.complex64)
