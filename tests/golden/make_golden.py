"""Writes tests/golden/fft_golden.npz: seeded inputs and their FP64 DFTs.

Inputs follow the workload definitions (U[-1,1) re/im from numpy
default_rng(seed), complex64); outputs are FP64 (numpy pocketfft in
complex128, itself pinned against the brute-force DFT in tests).  The
reference (Swift/Metal/vDSP) stores no vectors and cannot run here, so
these are the build's own fixtures.

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")


def uniform_c64(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=tuple(shape) + (2,))
    return x.astype(np.float32).view(np.complex64)[..., 0]


def main():
    d = {}
    for n in (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384):
        x = uniform_c64((1, n), 100 + n.bit_length())
        d[f"x_{n}"] = x[0]
        d[f"X_{n}"] = np.fft.fft(x[0].astype(np.complex128))
    # N=4096 batch 256 seed 0 (the plain forward workload): per-transform
    # FP64 energy sum|X|^2 and the sum of all bins, full output not stored.
    x = uniform_c64((256, 4096), 0)
    X = np.fft.fft(x.astype(np.complex128), axis=-1)
    d["cfg1_energy"] = np.sum(np.abs(X) ** 2, axis=-1)
    d["cfg1_binsum"] = np.sum(X, axis=-1)
    d["cfg1_first_row"] = X[0]
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
