"""Ragged batches (lengths 64 ... 16384 mixed) against FP64 per transform,
in every launch structure: one launch of length-sorted tickets (default),
one launch per length, one persistent launch fed by a device ticket counter
(work queue), and the split form (one launch per length of 4096 and longer
plus one persistent launch of short tiles); optionally replayed from a
hipGraph.  The ticket forms (default, per size, queue) run the same code per
ticket and give the same bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ragged_lengths(count, seed):
    # log-uniform over {64, ..., 16384}
    return (2 ** np.random.default_rng(seed).integers(6, 15, count)).astype(np.int64)


def ragged_data(lengths, seed):
    rng = np.random.default_rng(seed + 1000)
    tot = int(lengths.sum())
    return (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)


def check_all(oracle, lengths, x, y, inverse=False, tol=1e-6, stride=1):
    pos = 0
    worst = 0.0
    for i, n in enumerate(lengths):
        n = int(n)
        if i % stride == 0:
            seg = x[pos:pos + n]
            ref = oracle.f64_ifft(seg) if inverse else oracle.f64_fft(seg)
            worst = max(worst, oracle.rel_l2(y[pos:pos + n], ref))
        pos += n
    assert worst < tol, worst


MODES = [0, 1, 2, 3, 4, 6, 8, 10]


@pytest.mark.parametrize("mode", MODES)
def test_ragged_small(gpu, oracle, mode):
    import torch
    lengths = ragged_lengths(300, 1)
    x = ragged_data(lengths, 1)
    rp = gpu.RaggedFFT(lengths)
    assert rp.extent == x.size and rp.elements == x.size
    split = len({int(n) if n >= 4096 else 0 for n in lengths})  # one launch per length >= 4096
    expect = (len(set(lengths.tolist())) if mode & gpu.RAGGED_PER_SIZE else
              split if mode & gpu.RAGGED_SPLIT else 1)
    assert rp.launches(mode) == expect
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    rp.execute(xd, yd, mode=mode)
    torch.cuda.synchronize()
    check_all(oracle, lengths, x, yd.cpu().numpy())
    zd = torch.empty_like(xd)
    rp.execute(yd, zd, direction=gpu.INVERSE, mode=mode)
    torch.cuda.synchronize()
    assert oracle.rel_l2(zd.cpu().numpy(), x) < 1e-6


@pytest.mark.parametrize("lengths", [[64, 4096, 128, 1024] * 40, [8192, 64] * 30, [16384] * 5,
                                     [16384, 8192] * 7, [2048] * 3 + [8192]])
def test_ragged_tile_classes(gpu, oracle, lengths):
    """Plans with lengths of 4096 and longer, shorter lengths, or both, in
    every launch structure."""
    import torch
    lengths = np.array(lengths, np.int64)
    x = ragged_data(lengths, 5)
    rp = gpu.RaggedFFT(lengths)
    split = len({int(n) if n >= 4096 else 0 for n in lengths})  # one launch per length >= 4096
    assert rp.launches() == 1 and rp.launches(gpu.RAGGED_SPLIT) == split
    xd = torch.from_numpy(x).cuda()
    for mode in (gpu.RAGGED_DEFAULT, gpu.RAGGED_PER_SIZE, gpu.RAGGED_QUEUE, gpu.RAGGED_SPLIT):
        yd = torch.empty_like(xd)
        rp.execute(xd, yd, mode=mode)
        torch.cuda.synchronize()
        check_all(oracle, lengths, x, yd.cpu().numpy())


def test_ragged_matches_uniform_plans(gpu, oracle):
    """Each ragged transform equals the same-length plan's output to FP32."""
    import torch
    lengths = np.array([64, 16384, 128, 4096, 64, 8192, 256, 512, 1024, 2048] * 3, np.int64)
    x = ragged_data(lengths, 2)
    rp = gpu.RaggedFFT(lengths)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    rp.execute(xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    pos = 0
    for n in lengths:
        n = int(n)
        assert oracle.rel_l2(y[pos:pos + n], gpu.MetalFFT(n).forward(x[pos:pos + n])) < 1e-6
        pos += n


def test_ragged_custom_offsets_in_place(gpu, oracle):
    import torch
    lengths = np.array([4096, 64, 16384, 64], np.int64)
    offsets = np.array([100, 0, 5000, 30000], np.int64)  # gaps, arbitrary order
    rp = gpu.RaggedFFT(lengths, offsets)
    assert rp.extent == 30064
    buf = (np.random.default_rng(3).standard_normal(rp.extent) * (1 + 1j)).astype(np.complex64)
    d = torch.from_numpy(buf.copy()).cuda()
    rp.execute(d, d)  # in place
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    for n, o in zip(lengths, offsets):
        assert oracle.rel_l2(out[o:o + n], oracle.f64_fft(buf[o:o + n])) < 1e-6
    # untouched gap keeps its values
    assert np.array_equal(out[64:100], buf[64:100])


def test_graph_replays_match_eager(gpu, oracle):
    """Graph replays, interleaved with eager work and syncs, all equal the
    eager result."""
    import torch
    lengths = ragged_lengths(500, 4)
    x = ragged_data(lengths, 4)
    rp = gpu.RaggedFFT(lengths)
    xd = torch.from_numpy(x).cuda()
    eager = torch.empty_like(xd)
    rp.execute(xd, eager)
    torch.cuda.synchronize()
    g = torch.empty_like(xd)
    for r in range(4):
        g.fill_(float("nan"))
        rp.execute(xd, g, mode=gpu.RAGGED_GRAPH)
        torch.cuda.synchronize()
        assert torch.equal(g, eager), r
        _ = (xd * 2).sum().item()  # unrelated eager work between replays


def test_graph_cache_alternating_keys(gpu, oracle):
    """Forward x -> y and inverse y -> z alternate in graph mode (two cached
    graphs), then six buffer/direction keys cycle through the four slots
    (eviction and re-capture): every replay equals the eager result."""
    import torch
    lengths = ragged_lengths(400, 9)
    x = ragged_data(lengths, 9)
    rp = gpu.RaggedFFT(lengths)
    xd = torch.from_numpy(x).cuda()
    ref_y = torch.empty_like(xd)
    rp.execute(xd, ref_y)
    ref_z = torch.empty_like(xd)
    rp.execute(ref_y, ref_z, direction=gpu.INVERSE)
    torch.cuda.synchronize()
    y = torch.empty_like(xd)
    z = torch.empty_like(xd)
    for r in range(3):
        y.fill_(float("nan"))
        z.fill_(float("nan"))
        rp.execute(xd, y, mode=gpu.RAGGED_GRAPH)
        rp.execute(y, z, direction=gpu.INVERSE, mode=gpu.RAGGED_GRAPH)
        torch.cuda.synchronize()
        assert torch.equal(y, ref_y) and torch.equal(z, ref_z), r
    outs = [torch.empty_like(xd) for _ in range(6)]
    for r in range(2):
        for k, o in enumerate(outs):
            o.fill_(float("nan"))
            rp.execute(xd, o, mode=gpu.RAGGED_GRAPH | (gpu.RAGGED_QUEUE if k % 2 else 0))
        torch.cuda.synchronize()
        for k, o in enumerate(outs):
            assert torch.equal(o, ref_y), (r, k)
    assert oracle.rel_l2(ref_z.cpu().numpy(), x) < 1e-6


@pytest.mark.parametrize("lengths", [[64, 4096, 128, 1024] * 40, [8192, 64] * 31, [16384] * 5,
                                     [16384, 8192] * 7 + [8192], [2048] * 3 + [8192], [64] * 129,
                                     [4096] * 3, [256, 512, 1024, 2048] * 50])
def test_ticket_forms_agree(gpu, oracle, lengths):
    """The default ticket launch, one launch per length and the work queue run
    the same code per ticket: equal bit for bit, including partial tickets (an
    odd number of 8192 or 4096 transforms, a short last tile); the split form
    within FP32 rounding.  The queue counters reset themselves: repeated
    launches give the same bits."""
    import torch
    lengths = np.array(lengths, np.int64)
    x = ragged_data(lengths, 6)
    rp = gpu.RaggedFFT(lengths)
    assert rp.launches(gpu.RAGGED_QUEUE) == 1
    xd = torch.from_numpy(x).cuda()
    ref = torch.empty_like(xd)
    rp.execute(xd, ref)
    per = torch.empty_like(xd)
    rp.execute(xd, per, mode=gpu.RAGGED_PER_SIZE)
    spl = torch.empty_like(xd)
    rp.execute(xd, spl, mode=gpu.RAGGED_SPLIT)
    torch.cuda.synchronize()
    assert torch.equal(per, ref)
    assert oracle.rel_l2(spl.cpu().numpy(), ref.cpu().numpy()) < 1e-6
    q = torch.empty_like(xd)
    for rep in range(3):
        q.fill_(float("nan"))
        rp.execute(xd, q, mode=gpu.RAGGED_QUEUE)
        torch.cuda.synchronize()
        assert torch.equal(q, ref), rep
    check_all(oracle, lengths, x, ref.cpu().numpy())
    back = torch.empty_like(xd)
    rp.execute(ref, back, direction=gpu.INVERSE)
    torch.cuda.synchronize()
    assert oracle.rel_l2(back.cpu().numpy(), x) < 1e-6


def test_queue_graph_and_in_place(gpu, oracle):
    import torch
    lengths = np.array([4096, 64, 16384, 64, 8192], np.int64)
    offsets = np.array([100, 0, 5000, 30000, 40000], np.int64)
    rp = gpu.RaggedFFT(lengths, offsets)
    buf = (np.random.default_rng(8).standard_normal(rp.extent) * (1 - 1j)).astype(np.complex64)
    d = torch.from_numpy(buf.copy()).cuda()
    rp.execute(d, d, mode=gpu.RAGGED_QUEUE)  # in place
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    for n, o in zip(lengths, offsets):
        assert oracle.rel_l2(out[o:o + n], oracle.f64_fft(buf[o:o + n])) < 1e-6
    assert np.array_equal(out[64:100], buf[64:100])
    src = torch.from_numpy(buf).cuda()
    eager = torch.empty_like(src)
    rp.execute(src, eager, mode=gpu.RAGGED_QUEUE)
    g = torch.empty_like(src)
    for r in range(4):
        g.fill_(float("nan"))
        rp.execute(src, g, mode=gpu.RAGGED_QUEUE | gpu.RAGGED_GRAPH)
        torch.cuda.synchronize()
        for n, o in zip(lengths, offsets):
            assert torch.equal(g[o:o + n], eager[o:o + n]), r
        _ = (src * 2).sum().item()
    with pytest.raises(gpu.FFTError):
        rp.execute(src, g, mode=gpu.RAGGED_QUEUE | gpu.RAGGED_PER_SIZE)


def test_config2_ten_thousand(gpu, oracle):
    """10,000 transforms, lengths log-uniform over {64..16384} (seed 1), N(0,1)
    values: forward + inverse in one launch, against the other launch
    structures; FP64 on a sample."""
    import torch
    lengths = ragged_lengths(10000, 1)
    rng = np.random.default_rng(1)
    tot = int(lengths.sum())
    x = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    rp = gpu.RaggedFFT(lengths)
    xd = torch.from_numpy(x).cuda()
    y1 = torch.empty_like(xd)
    y2 = torch.empty_like(xd)
    y3 = torch.empty_like(xd)
    y4 = torch.empty_like(xd)
    assert rp.launches() == 1
    rp.execute(xd, y1)
    rp.execute(xd, y2, mode=gpu.RAGGED_PER_SIZE)
    rp.execute(xd, y3, mode=gpu.RAGGED_QUEUE)
    rp.execute(xd, y4, mode=gpu.RAGGED_SPLIT | gpu.RAGGED_GRAPH)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)  # same math, different launch structure
    assert torch.equal(y1, y3)
    assert oracle.rel_l2(y4.cpu().numpy(), y1.cpu().numpy()) < 1e-6
    check_all(oracle, lengths, x, y1.cpu().numpy(), stride=97)
    z = torch.empty_like(xd)
    rp.execute(y1, z, direction=gpu.INVERSE)
    torch.cuda.synchronize()
    assert oracle.rel_l2(z.cpu().numpy(), x) < 1e-6


def test_ragged_errors(gpu):
    with pytest.raises(gpu.FFTError) as ei:
        gpu.RaggedFFT([64, 128, 100])
    assert ei.value.case == "unsupportedFFTSize" and ei.value.batch_index == 2
    rp = gpu.RaggedFFT([])
    assert rp.launches() == 0
