"""The C++ mirror of the Swift API (metalfft_amd/host/metalfft.hpp) through
its self-test executable, built in-tree with the library."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_host_mirror_selftest():
    exe = os.path.join(ROOT, "metalfft_amd", "mifft_host_selftest")
    assert os.path.exists(exe), "build first: python -m metalfft_amd.build"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
