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
