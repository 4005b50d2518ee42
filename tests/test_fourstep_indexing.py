"""CPU emulation of the four-step kernels' index arithmetic (fourstep.hip).

For every compiled four-step variant (names from the library's variant
list), the grid, the staging of C-column tiles (element (c, r) at
src[c + r * ld], lanes tid + THREADS q), the W_N^{n1 k2} factor from the two
tables (hi[e >> S] * lo[e & (2^S - 1)]) and the transposed store are
replayed in numpy with exact sub-transforms: every load and store must stay
inside its transform, every intermediate and output element must be written
exactly once, and the result must be the length-N DFT."""
import re

import numpy as np
import pytest

import metalfft_amd as m

FOURSTEP_SIZES = (32768, 65536, 131072, 262144, 524288, 1 << 20)


def variants():
    out = []
    for n in FOURSTEP_SIZES:
        for name in m.kernel_variants(n):
            g = re.fullmatch(r"fourstep_(\d+)x(\d+)_c(\d+)_c(\d+)", name)
            if g:
                out.append((n, *map(int, g.groups())))
    return out


def emulate(x, n1s, n2s, ca, cb):
    batch, n = x.shape
    assert n == n1s * n2s
    s = (n.bit_length() - 1 + 1) // 2
    lo = np.exp(-2j * np.pi * np.arange(1 << s) / n)
    hi = np.exp(-2j * np.pi * (np.arange(n >> s) << s) / n)
    xf = x.reshape(-1)
    y = np.zeros(batch * n, np.complex128)
    ycount = np.zeros(batch * n, np.int64)
    out = np.zeros(batch * n, np.complex128)
    ocount = np.zeros(batch * n, np.int64)

    # pass A: CA columns of N2 rows, stride N1
    threads = ca * (n2s // 16)
    f = np.arange(threads * 16)
    c, r = f % ca, f // ca
    for blk in range(batch * (n1s // ca)):
        tr, c0 = blk // (n1s // ca), (blk % (n1s // ca)) * ca
        idx = tr * n + c0 + c + r * n1s
        assert idx.min() >= tr * n and idx.max() < (tr + 1) * n
        tile = np.zeros((ca, n2s), np.complex128)
        seen = np.zeros((ca, n2s), np.int64)
        tile[c, r] = xf[idx]
        np.add.at(seen, (c, r), 1)
        assert (seen == 1).all()
        col = np.fft.fft(tile, axis=1)
        k2 = np.arange(n2s)
        for local in range(ca):
            n1 = c0 + local
            e = n1 * k2
            assert e.max() < n
            w = hi[e >> s] * lo[e & ((1 << s) - 1)]
            dst = tr * n + n1 * n2s + k2
            y[dst] = col[local] * w
            ycount[dst] += 1
    assert (ycount == 1).all()

    # pass B: CB columns of N1 rows of y, stride N2; transposed store
    threads = cb * (n1s // 16)
    f = np.arange(threads * 16)
    c, r = f % cb, f // cb
    for blk in range(batch * (n2s // cb)):
        tr, k0 = blk // (n2s // cb), (blk % (n2s // cb)) * cb
        idx = tr * n + k0 + c + r * n2s
        assert idx.min() >= tr * n and idx.max() < (tr + 1) * n
        tile = np.zeros((cb, n1s), np.complex128)
        tile[c, r] = y[idx]
        col = np.fft.fft(tile, axis=1)
        out[idx] = col[c, r]  # stage_out reads (c, r) and writes the same address set
        ocount[idx] += 1
    assert (ocount == 1).all()
    return out.reshape(batch, n)


def test_variant_list_covers_every_length():
    got = {v[0] for v in variants()}
    assert got == set(FOURSTEP_SIZES)


@pytest.mark.parametrize("n,n1,n2,ca,cb", variants())
def test_fourstep_indexing(n, n1, n2, ca, cb):
    assert n1 * n2 == n and n1 % ca == 0 and n2 % cb == 0
    assert ca * (n2 // 16) <= 1024 and cb * (n1 // 16) <= 1024
    pad = lambda c: 1 if c >= 32 else 32 // c  # noqa: E731
    assert ca * (n2 + pad(ca)) * 8 <= 160 * 1024 and cb * (n1 + pad(cb)) * 8 <= 160 * 1024
    rng = np.random.default_rng(n + ca)
    batch = 2 if n <= 131072 else 1
    x = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    got = emulate(x, n1, n2, ca, cb)
    ref = np.fft.fft(x, axis=1)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
