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
