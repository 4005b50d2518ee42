"""Build recipe for libmifft.so: every HIP source under csrc/ compiled for
gfx950 with hipcc, linked in-tree into metalfft_amd/libmifft.so.

Run ``python -m metalfft_amd.build`` (or ``__graft_entry__.build()``).  It
always recompiles; objects go to metalfft_amd/_build/ (git-ignored).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmifft.so")
SELFTEST = os.path.join(HERE, "mifft_host_selftest")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ mirror self-test, linked against the library (rpath: its directory)
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "host", "selftest.cpp"), "-o", SELFTEST + ".tmp",
           "-L", HERE, "-lmifft", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"self-test build failed:\n{r.stderr}")
    os.replace(SELFTEST + ".tmp", SELFTEST)
    if verbose:
        print(f"built {LIB} from {len(objs)} sources (+ {os.path.basename(SELFTEST)})",
              file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
