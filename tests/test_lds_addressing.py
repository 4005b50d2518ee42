"""CPU checks of the index identities the LDS exchange code relies on
(metalfft_amd/csrc/stockham_core.hpp: LdsSwz, lds_read_column,
lds_write_butterfly, the split-LDS exchange and the radix schedule), over
every length, values-per-thread count and pass the kernels instantiate."""
import pytest

SIZES = [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]


def swz(i):
    """LdsSwz::idx: flip bits 0-4 by bits 4-8."""
    return i ^ ((i >> 4) & 31)


def schedule(n, e):
    """Schedule<N, E>: the remainder radix first, then radix min(E, 16)."""
    rm = min(e, 16)
    lg, lr = n.bit_length() - 1, rm.bit_length() - 1
    full, rem = divmod(lg, lr)
    radix = ([1 << rem] if rem else []) + [rm] * full
    stride = [1]
    for r in radix[:-1]:
        stride.append(stride[-1] * r)
    return radix, stride


@pytest.mark.parametrize("n", SIZES)
def test_schedule_covers_n(n):
    for e in (4, 8, 16, 32):
        if e > n:
            continue
        radix, stride = schedule(n, e)
        prod = 1
        for r in radix:
            prod *= r
        assert prod == n
        assert radix[-1] * stride[-1] == n  # last pass leaves v[m] = X[t + T m]
        assert all(r <= 16 for r in radix) and all(r == radix[-1] for r in radix[1:])


@pytest.mark.parametrize("n", SIZES)
def test_swizzle_is_a_bijection(n):
    assert sorted(swz(i) for i in range(n)) == list(range(n))
    if n >= 1024:  # the split-LDS buffers hold N/2 entries
        assert sorted(swz(i) for i in range(n // 2)) == list(range(n // 2))


@pytest.mark.parametrize("t_count", [32, 64, 128, 256, 512, 1024, 2048])
def test_column_read_identity(t_count):
    """idx(t + T m) = (idx(t) ^ c(m)) + T m, c(m) = ((T/16) m % P) & 31."""
    period = 1 if t_count >= 512 else 512 // t_count
    for t in range(t_count):
        i0 = swz(t)
        for m in range(32):
            c = ((t_count // 16) * (m % period)) & 31
            assert swz(t + t_count * m) == (i0 ^ c) + t_count * m


@pytest.mark.parametrize("n", SIZES)
def test_butterfly_write_identity(n):
    """idx(base + j s) = idx(base) ^ idx(j s) for every butterfly base of
    every pass that writes to LDS, and in the split exchange for the bases
    shifted down by a multiple of N/Q."""
    for e in (4, 8, 16, 32):
        if e > n:
            continue
        radix, stride = schedule(n, e)
        for r, s in zip(radix[:-1], stride[:-1]):
            for b in range(n // r):
                base = (b // s) * (r * s) + (b % s)
                i0 = swz(base)
                for j in range(r):
                    assert swz(base + j * s) == i0 ^ swz(j * s)
                if e == 32:
                    h = 1 if base >= n // 2 else 0
                    sb = base - h * (n // 2)
                    for j in range(r):
                        assert swz(sb + j * s) == swz(sb) ^ swz(j * s)


@pytest.mark.parametrize("n", [2048, 4096, 8192, 16384, 32768])
def test_split_exchange_halves(n):
    """E = 32: butterfly i of every thread writes half i // (B/2), each half
    exactly once, and half h is read back as v[m], m in [16h, 16h + 16)."""
    e, t_count = 32, n // 32
    radix, stride = schedule(n, e)
    for r, s in zip(radix[:-1], stride[:-1]):
        bpt = e // r
        hits = [0] * n
        for t in range(t_count):
            for i in range(bpt):
                b = t + i * t_count
                base = (b // s) * (r * s) + (b % s)
                h = i // (bpt // 2)
                for j in range(r):
                    p = base + j * s
                    assert p // (n // 2) == h
                    hits[p] += 1
        assert hits == [1] * n
    for t in range(t_count):
        for m in range(e):
            assert (t + t_count * m) // (n // 2) == m // 16


@pytest.mark.parametrize("n", SIZES)
def test_split_exchange_halves_e16(n):
    """E = 16 (k_stockham_half16, and at every length for shorter tiles): passes with several
    butterflies per thread split them as for E = 32; radix-16 passes (one
    butterfly per thread) split by thread, half 0 written by threads t < T/2
    (whole waves from 8192 up), each half exactly once; half h is read back
    as v[m], m in [8h, 8h + 8); the XOR write identity holds for the
    shifted bases."""
    e, t_count = 16, n // 16
    radix, stride = schedule(n, e)
    for r, s in zip(radix[:-1], stride[:-1]):
        bpt = e // r
        hits = [0] * n
        for t in range(t_count):
            for i in range(bpt):
                b = t + i * t_count
                base = (b // s) * (r * s) + (b % s)
                h = i // (bpt // 2) if bpt >= 2 else (0 if t < t_count // 2 else 1)
                if bpt == 1:
                    assert (t < t_count // 2) == (b < n // (2 * r))
                    if n >= 8192:
                        assert (t_count // 2) % 64 == 0  # whole waves on either side
                sb = base - h * (n // 2)
                for j in range(r):
                    p = base + j * s
                    assert p // (n // 2) == h
                    assert swz(sb + j * s) == swz(sb) ^ swz(j * s)
                    hits[p] += 1
        assert hits == [1] * n
    for t in range(t_count):
        for m in range(e):
            assert (t + t_count * m) // (n // 2) == m // 8


def test_mfma_pass_addresses():
    """The DFT-as-GEMM kernel's closed-form LDS addresses
    (metalfft_amd/csrc/mfma_fft.hip) equal midx() of the Stockham positions
    for every lane (w, c, g), tile t4 and element i."""
    def midx(i):
        return i + (i >> 4) + ((i >> 10) << 4)
    for w in range(4):
        for t4 in range(4):
            for c in range(16):
                for g in range(4):
                    for i in range(4):
                        b, j = 16 * (4 * w + t4) + c, 4 * g + i
                        assert midx(b + 256 * j) == 68 * w + c + 1104 * g + 17 * t4 + 272 * i
                        assert midx(16 * b + j) == 1104 * w + 17 * c + 4 * g + 272 * t4 + i
                        assert midx(256 * (b // 16) + b % 16 + 16 * j) == \
                            1104 * w + c + 68 * g + 272 * t4 + 17 * i


def test_mfma_radix64_mapping():
    """The radix-64 MFMA form (metalfft_amd/csrc/mfma_fft.hip, R64): pass 0
    reads X_J of butterfly B = 16 w + c at B + 64 J, J = t4 + 16 g + 4 i
    (tile t4 = j4, MFMA column c, K range j16 = 4 g + i), four DFT-16s give
    Z_t4[k16] (k16 = 4 g + i in the lane's D rows), the twiddle W_64^{t4 k16}
    and a DFT-4 over t4 give Y[k16 + 16 k4], written to 64 B + K; pass 1
    reads the same positions with W_4096^{J B} and writes B + 64 K.  Run in
    FP64 with exactly these indices it is the DFT; r64idx is a bijection and
    every half-wave of either access lands on 32 distinct 8-byte slots."""
    import numpy as np
    n = 4096

    def r64idx(a):
        m = ((a >> 6) & 3) | (((a >> 8) & 1) << 3) | ((((a >> 9) ^ (a >> 10)) & 1) << 4)
        return a ^ m

    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    f16 = np.exp(-2j * np.pi * np.outer(np.arange(16), np.arange(16)) / 16)
    rng = np.random.default_rng(64)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    lds = np.zeros(n, complex)
    out = np.zeros(n, complex)
    for p in range(2):
        new = np.zeros(n, complex)
        for w in range(4):
            for c in range(16):
                b = 16 * w + c
                z = np.zeros((4, 16), complex)
                for t4 in range(4):
                    xs = np.zeros(16, complex)
                    for g in range(4):
                        for i in range(4):
                            j = t4 + 16 * g + 4 * i
                            if p == 0:
                                xs[4 * g + i] = x[16 * w + c + 64 * t4 + 1024 * g + 256 * i]
                            else:
                                wj = tw[b * (16 * g + 4 * i)] * tw[b * t4]
                                xs[4 * g + i] = lds[r64idx(16 * w + c + 64 * j)] * wj
                    z[t4] = f16 @ xs
                for k16 in range(16):
                    w1 = tw[64 * k16]
                    a0, a1, a2, a3 = z[0, k16], z[1, k16] * w1, z[2, k16] * w1 * w1, z[3, k16] * w1 ** 3
                    y = [a0 + a1 + a2 + a3, (a0 - a2) - 1j * (a1 - a3), (a0 + a2) - (a1 + a3),
                         (a0 - a2) + 1j * (a1 - a3)]
                    for k4 in range(4):
                        kk = k16 + 16 * k4
                        if p == 0:
                            new[r64idx(64 * b + kk)] = y[k4]
                        else:
                            out[b + 64 * kk] = y[k4]
        lds = new
    ref = np.fft.fft(x)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-13
    a = np.arange(n)
    assert sorted(r64idx(a).tolist()) == list(range(n))
    for w in range(4):
        for h in range(2):
            half = [(c, g) for c in range(16) for g in (2 * h, 2 * h + 1)]
            for i in range(4):
                for k4 in range(4):
                    slots = {r64idx(64 * (16 * w + c) + 4 * g + i + 16 * k4) % 32 for c, g in half}
                    assert len(slots) == 32
                for t4 in range(4):
                    slots = {r64idx(16 * w + c + 64 * (t4 + 16 * g + 4 * i)) % 32 for c, g in half}
                    assert len(slots) == 32


def test_padded_layout_identities():
    """LdsPad (one spare entry per 16): idx(t + T m) = idx(t) + (T + T/16) m
    for T a multiple of 16, and idx(base + j s) = idx(base) + j (s + s/16)
    for s a multiple of 16, as lds_read_column / lds_write_butterfly use."""
    def pad(i):
        return i + (i >> 4)
    for t_count in (16, 32, 64, 128, 256, 512, 1024):
        for t in range(t_count):
            for m in range(16):
                assert pad(t + t_count * m) == pad(t) + (t_count + t_count // 16) * m
    for n in SIZES:
        for e in (8, 16, 32):
            if e > n:
                continue
            radix, stride = schedule(n, e)
            for r, s in zip(radix[:-1], stride[:-1]):
                if s % 16:
                    continue
                for b in range(n // r):
                    base = (b // s) * (r * s) + (b % s)
                    for j in range(r):
                        assert pad(base + j * s) == pad(base) + j * (s + s // 16)


def test_corner_turn_tile_swizzle():
    """Complex-FP16 corner turn tile (metalfft_amd/csrc/sar.hip,
    k_corner_turn_h16): element (r, c) at word r*64 + (c ^ 4*((r >> 2) & 15))
    is a bijection, keeps each row's 16-byte chunks whole, and the
    transposed reads of one wave hit 64 distinct banks."""
    def word(r, c):
        return r * 64 + (c ^ (4 * ((r >> 2) & 15)))
    assert sorted(word(r, c) for r in range(64) for c in range(64)) == list(range(4096))
    for r in range(64):
        for q in range(16):
            for k in range(4):
                assert word(r, 4 * q + k) == r * 64 + 4 * (q ^ ((r >> 2) & 15)) + k
    for w in range(4):
        for p in range(4):
            for i in range(4):
                banks = set()
                for lane in range(64):
                    tid = 64 * w + lane
                    q, c = tid & 15, (tid >> 4) + 16 * p
                    assert word(4 * q + i, c) == (4 * q + i) * 64 + (c ^ (4 * q))
                    banks.add(word(4 * q + i, c) % 64)
                assert len(banks) == 64
