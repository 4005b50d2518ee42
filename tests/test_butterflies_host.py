"""The in-register DFT butterflies of the HIP kernels (fft_device.hpp, incl.
the twiddle-folded radix-16 form) compiled for the host and checked against
an FP64 DFT of the same FP32 inputs (tests/host/butterfly_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc absent")
def test_butterflies_match_fp64_dft(tmp_path):
    exe = str(tmp_path / "butterfly_check")
    src = os.path.join(HERE, "host", "butterfly_check.cpp")
    inc = os.path.join(ROOT, "metalfft_amd", "csrc")
    cc = HIPCC if os.path.exists(HIPCC) else "hipcc"
    r = subprocess.run([cc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", inc, src, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
