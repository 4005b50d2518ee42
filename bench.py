"""Benchmark: batched complex FFT N=4096, batch 4096, complex64, on one GPU.

One step = one batched forward transform of 4096 length-4096 complex64
signals (128 MiB in, 128 MiB out) already resident in HBM.  Input sets
rotate over >= 1 GiB so the 256 MiB Infinity Cache cannot hold a step's
working set.  Prints one JSON line (metric, value, timing, configuration);
extra keys carry the roofline, the CPU baseline and the reduced-precision
forms.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20]
"""
from __future__ import annotations

import argparse
import contextlib
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
METRIC = "complex FFTs/s + % HBM roofline, N=4096 b=4096; FP16-MFMA vs FP32 at stated err"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=4096)
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--sets", type=int, default=4, help="rotating in/out buffer sets")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--launch", choices=("graph", "stream"), default="graph",
                   help="timed region: one hipGraph replay of the K launches (default) or K "
                        "launches from a Python loop; the other is reported beside it")
    p.add_argument("--kernel", default=None, help="pin a kernel variant by name")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the MFMA / ragged / convolve / SAR measurements")
    p.add_argument("--only", default=None,
                   help="comma-separated extras to run (default: all), e.g. stft_welch_2p24")
    p.add_argument("--replays", type=int, default=31,
                   help="timed graph replays of the K launches (back to back); the median "
                        "is reported")
    p.add_argument("--settle", type=float, default=3.0,
                   help="at most this many seconds of untimed replays before timing, until "
                        "the replay time stops falling")
    p.add_argument("--probe", default=None, choices=sorted(PROBES),
                   help="only run a few launches of one kernel's workload (for rocprofv3 "
                        "--pmc counter runs) and exit")
    p.add_argument("--traffic-probe", action="store_true",
                   help="only run 4 launches of a plain 128 MiB copy (elementwise, known "
                        "bytes) and 4 of the FFT step, for rocprofv3 --pmc FETCH_SIZE / "
                        "WRITE_SIZE runs")
    return p.parse_args()


@contextlib.contextmanager
def no_gc():
    """Python's cyclic collector off around a timed launch loop: a
    collection that falls inside it (after the large allocations of earlier
    measurements) stalls the host for a millisecond or more between two
    launches, which the events then count as GPU time.  No collection is
    forced here: the milliseconds of host work it takes would leave the GPU
    idle right before the timed launches, and its clocks drop meanwhile
    (measured: 3-4 us more per 4096 x 4096 batch)."""
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def time_gpu(fn, steps, warmup, torch):
    """Stream loop: `steps` back-to-back launches from Python, HIP events on
    the launch stream around them."""
    for i in range(warmup):
        fn(i, None)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    with no_gc():
        t0 = time.perf_counter()
        e0.record(stream)
        for i in range(steps):
            fn(i, None)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def time_graph(fn, steps, warmup, torch, reps=1, settle_s=0.0):
    """The same `steps` launches captured once into a hipGraph and replayed:
    the GPU runs them back to back with no host launch jitter between them.
    Warm-up launches run eagerly first (plans and occupancy queries settle
    outside the capture).  With settle_s > 0, untimed replays then run back
    to back in chunks of about 2 ms of GPU time until the per-replay time of
    a chunk stops falling (no new minimum, by 0.3 %, in three chunks) or
    settle_s seconds have passed: the clocks reach the state a long run
    holds (a 20-launch replay right after start-up reads several percent
    slow).  Then `reps` replays are enqueued back to back, each bracketed by
    its own pair of HIP events on the replay stream, with one
    synchronisation after the last: the result is the median replay time
    per launch.  Returns (ms per launch, wall seconds of the timed replays,
    graph, [ms per launch of each timed replay], settle replays run)."""
    for i in range(warmup):
        fn(i, None)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        capture = torch.cuda.current_stream().cuda_stream
        for i in range(steps):
            fn(i, capture)
    stream = torch.cuda.current_stream()

    def timed(count):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(count)]
        for e0, e1 in ev:
            e0.record(stream)
            graph.replay()
            e1.record(stream)
        torch.cuda.synchronize()
        return [e0.elapsed_time(e1) for e0, e1 in ev]

    settled = 0
    with no_gc():
        one = min(timed(2))
        if settle_s > 0:
            chunk = max(1, int(2.0 / max(one, 1e-3)))  # ~2 ms of GPU time per chunk
            best, since, t_start = float("inf"), 0, time.perf_counter()
            while since < 3 and time.perf_counter() - t_start < settle_s:
                t = sorted(timed(chunk))[chunk // 2]
                settled += chunk
                if t < best * 0.997:
                    best, since = t, 0
                else:
                    since += 1
        t0 = time.perf_counter()
        times = timed(reps)
        wall = time.perf_counter() - t0
    per = sorted(t / steps for t in times)
    return per[len(per) // 2], wall, graph, [t / steps for t in times], settled


def copy_ceiling(sets, steps, torch):
    """Achievable HBM copy rate on this box: the library's diagnostic copy
    (mifft_copy_device: 16-byte non-temporal loads and stores, the FFT
    kernels' access style) over the same rotating 128 MiB sets, timed like
    the FFT (one hipGraph replay of `steps` launches)."""
    import ctypes
    from metalfft_amd import _lib
    L = _lib.load()
    nbytes = sets[0][0].numel() * sets[0][0].element_size()

    def fn(i, stream):
        src, dst = sets[i % len(sets)]
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _lib.check(L.mifft_copy_device(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                       ctypes.c_int64(nbytes), ctypes.c_void_p(s)))
    ms, _, graph, _, _ = time_graph(fn, steps, 10, torch, reps=3, settle_s=1.0)
    del graph
    return {"gbs": 2 * nbytes / (ms * 1e-3) / 1e9, "ms_per_copy": ms, "bytes_per_copy": 2 * nbytes,
            "kernel": "mifft_copy_device, 4 rotating 128 MiB sets",
            "method": "median of 3 back-to-back replays of the K launches after settling"}


def cpu_baseline(n, batch, torch_seed_data):
    """The oracle's FP32 restatement of the reference FFT on the host cores:
    reference protocol (5 warmups, median of 20), all cores; plus 1 core."""
    from oracle import oracle as O
    O.build()
    x = torch_seed_data
    threads = min(16, os.cpu_count() or 1)
    out = {}
    for label, th, warm, reps in (("all", threads, 5, 20), ("one", 1, 1, 5)):
        for _ in range(warm):
            O.cpu_fft(x, threads=th)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            O.cpu_fft(x, threads=th)
            ts.append(time.perf_counter() - t0)
        out[label] = (batch / float(np.median(ts)), th)
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                     if l.startswith("model name"))
    except Exception:
        model = "unknown"
    v, th = out["all"]
    return {"value": v, "unit": "transforms/s", "cores": th, "kind": "port",
            "sample": f"full workload ({batch} x N={n} complex64 U[-1,1) seed 0), "
                      f"median of 20 after 5 warmups, OpenMP over the batch",
            "one_core_value": out["one"][0], "cpu_model": model,
            "host_cpus": os.cpu_count()}


def rel_l2(a, b):
    """sqrt(sum |a - b|^2 / sum |b|^2) in FP64."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = float(np.sum(np.abs(b) ** 2))
    num = float(np.sum(np.abs(a - b) ** 2))
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.complex128) - np.asarray(b, np.complex128))))


def time_fn(fn, reps, warm, torch):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    with no_gc():
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tones_gpu(batch, n, seed, torch, tones=20, noise=1e-3):
    """Sum of 20 complex tones per transform (random frequency / phase,
    amplitudes log-uniform over 60 dB) plus complex N(0, noise^2), on the GPU."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    f = torch.rand((batch, tones, 1), generator=g, device="cuda", dtype=torch.float64) * n
    ph = torch.rand((batch, tones, 1), generator=g, device="cuda", dtype=torch.float64) * 2 * np.pi
    amp = 10.0 ** (-3.0 * torch.rand((batch, tones, 1), generator=g, device="cuda",
                                     dtype=torch.float64))
    k = torch.arange(n, device="cuda", dtype=torch.float64)
    x = torch.zeros((batch, n), dtype=torch.complex128, device="cuda")
    for t in range(tones):  # one tone at a time keeps memory at batch x n
        x += amp[:, t] * torch.exp(1j * (2 * np.pi * f[:, t] * k / n + ph[:, t]))
    x += noise * torch.complex(torch.randn((batch, n), generator=g, device="cuda", dtype=torch.float64),
                               torch.randn((batch, n), generator=g, device="cuda", dtype=torch.float64))
    return x


def extra_mfma(m, torch, n=4096, batch=4096):
    """FP16 / BF16 DFT-as-GEMM forms, plain and hi/lo split, complex64 and
    complex-FP16 device data, on the sum-of-tones workload; rel-L2 / max-abs
    vs FP64 on 64 sampled transforms."""
    x64 = tones_gpu(batch, n, 3, torch)
    idx = torch.arange(0, batch, batch // 64, device="cuda")
    ref = np.fft.fft(x64[idx].cpu().numpy(), axis=-1)
    xc = x64.to(torch.complex64)
    del x64
    xh = torch.view_as_real(xc).reshape(batch, 2 * n).half().contiguous()
    out = {"note": "4 rotating input/output sets per form (>= 512 MiB), bytes per I/O type"}
    rot = {0: [xc] + [xc.clone() for _ in range(3)], 1: [xh] + [xh.clone() for _ in range(3)]}
    outs = {io: [torch.empty_like(v[0]) for _ in range(4)] for io, v in rot.items()}

    def timed(plan, io):
        state = {"i": 0}

        def step():
            i = state["i"] % 4
            plan.execute(rot[io][i], outs[io][i], batch)
            state["i"] += 1
        return time_fn(step, 20, 4, torch)

    names = {1: "fp16", 2: "bf16", 3: "fp16_split", 4: "bf16_split"}
    for prec, name in names.items():
        for io in (0, 1):
            key = f"{name}_{'c64' if io == 0 else 'c32h'}"
            try:
                plan = m.PrecisionFFT(n, prec, io)
                ms = timed(plan, io)
                plan.execute(rot[io][0], outs[io][0], batch)
                torch.cuda.synchronize()
                dst = outs[io][0]
                if io == 0:
                    y = dst[idx].cpu().numpy()
                else:
                    y = dst[idx].float().cpu().numpy().view(np.complex64)
                nbytes = (16 if io == 0 else 8) * n * batch
                out[key] = {"transforms_per_s": batch / (ms * 1e-3), "ms": ms,
                            "gbs": nbytes / (ms * 1e-3) / 1e9,
                            "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            "rel_l2": rel_l2(y, ref),
                            "max_abs": max_abs(y, ref), "kernel": plan.kernel_name}
            except Exception as e:  # report, keep the other lines
                out[key] = {"error": repr(e)}
    # the FP32 path on the same data, for the same error metric
    fft = m.MetalFFT(n)
    state = {"i": 0}

    def step32():
        i = state["i"] % 4
        fft.execute(rot[0][i], outs[0][i], batch)
        state["i"] += 1
    ms = time_fn(step32, 20, 4, torch)
    fft.execute(rot[0][0], outs[0][0], batch)
    torch.cuda.synchronize()
    out["fp32_c64"] = {"transforms_per_s": batch / (ms * 1e-3), "ms": ms,
                       "gbs": 16 * n * batch / (ms * 1e-3) / 1e9,
                       "rel_l2": rel_l2(outs[0][0][idx].cpu().numpy(), ref),
                       "max_abs": max_abs(outs[0][0][idx].cpu().numpy(), ref),
                       "kernel": fft.kernel_name}
    return out


def extra_ragged(m, torch, count=10000, cpu=True):
    lengths = (2 ** np.random.default_rng(1).integers(6, 15, count)).astype(np.int64)
    rp = m.RaggedFFT(lengths)
    tot = int(lengths.sum())
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.complex(torch.randn(tot, generator=g, device="cuda"),
                      torch.randn(tot, generator=g, device="cuda"))
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    out = {"transforms": count, "elements": tot, "bytes_per_direction": 16 * tot}
    for name, mode in (("one_launch", m.RAGGED_DEFAULT),
                       ("one_launch_graph", m.RAGGED_DEFAULT | m.RAGGED_GRAPH),
                       ("work_queue", m.RAGGED_QUEUE), ("work_queue_graph", m.RAGGED_QUEUE | m.RAGGED_GRAPH),
                       ("split", m.RAGGED_SPLIT), ("split_graph", m.RAGGED_SPLIT | m.RAGGED_GRAPH),
                       ("per_size", m.RAGGED_PER_SIZE),
                       ("per_size_graph", m.RAGGED_PER_SIZE | m.RAGGED_GRAPH)):
        def step():
            rp.execute(x, y, m.FORWARD, mode)
            rp.execute(y, z, m.INVERSE, mode)
        ms = time_fn(step, 10, 2, torch)
        gbs = 2 * 16 * tot / (ms * 1e-3) / 1e9
        out[name] = {"ms_fwd_inv": ms, "transforms_per_s": 2 * count / (ms * 1e-3),
                     "gbs": gbs, "roofline_frac": gbs / HBM_PEAK_GBS, "launches": rp.launches(mode)}
    # accuracy of the default form: every 50th transform against numpy's FP64
    # FFT, and the forward + inverse round trip over all of them
    rp.execute(x, y, m.FORWARD, m.RAGGED_DEFAULT)
    rp.execute(y, z, m.INVERSE, m.RAGGED_DEFAULT)
    torch.cuda.synchronize()
    xh_, yh_ = x.cpu().numpy(), y.cpu().numpy()
    offs_ = np.concatenate([[0], np.cumsum(lengths)])
    got, ref = [], []
    for i in range(0, count, 50):
        seg = slice(offs_[i], offs_[i + 1])
        got.append(yh_[seg])
        ref.append(np.fft.fft(xh_[seg].astype(np.complex128)))
    got, ref = np.concatenate(got), np.concatenate(ref)
    out["accuracy"] = {"rel_l2_vs_fp64": rel_l2(got, ref), "max_abs_vs_fp64": max_abs(got, ref),
                       "sample": f"{len(range(0, count, 50))} transforms (every 50th), all lengths",
                       "roundtrip_rel_l2": rel_l2(z.cpu().numpy(), xh_)}
    if cpu:
        # the oracle's FP32 restatement over the same lengths, grouped by size
        from oracle import oracle as O
        O.build()
        xh = x.cpu().numpy()
        offs = np.concatenate([[0], np.cumsum(lengths)])
        groups = {}
        for i, n in enumerate(lengths):
            groups.setdefault(int(n), []).append(xh[offs[i]:offs[i + 1]])
        groups = {n: np.stack(v) for n, v in groups.items()}
        threads = min(16, os.cpu_count() or 1)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            for n, v in groups.items():
                O.cpu_fft(O.cpu_fft(v, threads=threads), inverse=True, threads=threads)
            ts.append(time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": 2 * count / float(np.median(ts)), "unit": "transforms/s",
                               "cores": threads, "kind": "port",
                               "sample": "the same 10,000 transforms forward + inverse, median of 3"}
    return out


def extra_convolve(m, torch, count=512, S=65536, K=1023, N=4096, cpu=True):
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((count, S), generator=g, device="cuda")
    h = torch.randn(K, generator=g, device="cuda") / np.sqrt(K)
    cv = m.FusedConvolver(h.cpu().numpy(), N)
    y = torch.empty((count, S + K - 1), device="cuda")
    gy = torch.randn_like(y)
    gx = torch.empty_like(x)
    gh = torch.empty(K, device="cuda")
    fwd = time_fn(lambda: cv.apply_device(x, y, S, count), 10, 2, torch)
    bwd = time_fn(lambda: cv.backward_device(x, gy, gx, gh, S, count), 10, 2, torch)
    fbytes = 4 * count * (S + S + K - 1)
    bbytes = 4 * count * (S + (S + K - 1) + S)
    fp16 = {}
    for name in ("float16Pure", "float16Storage", "float16Mixed"):
        c16 = m.FusedConvolver(h.cpu().numpy(), N, getattr(m.Precision, name))
        ms = time_fn(lambda: c16.apply_device(x, y, S, count), 10, 2, torch)
        fp16[name] = {"forward_ms": ms, "forward_gbs_alg": fbytes / (ms * 1e-3) / 1e9}
        fp16[name]["_out"] = y[[0, count - 1]].cpu().numpy()
        c16.close()
    # accuracy against FP64 (scipy's FFT convolution in float64): the
    # forward of two signals in every precision, the adjoint's grad_x on the
    # same two and grad_h over all signals
    from scipy.signal import fftconvolve
    sig = [0, count - 1]
    xs64 = x[sig].cpu().numpy().astype(np.float64)
    h64 = h.cpu().numpy().astype(np.float64)
    y_ref = np.stack([fftconvolve(r, h64) for r in xs64])
    acc = {}
    cv.apply_device(x, y, S, count)
    torch.cuda.synchronize()
    acc["forward_fp32"] = {"rel_l2": rel_l2(y[sig].cpu().numpy(), y_ref),
                           "max_abs": max_abs(y[sig].cpu().numpy(), y_ref)}
    cv.backward_device(x, gy, gx, gh, S, count)
    torch.cuda.synchronize()
    gy64 = gy[sig].cpu().numpy().astype(np.float64)
    gx_ref = np.stack([fftconvolve(r, h64[::-1], mode="valid") for r in gy64])
    acc["grad_signal"] = {"rel_l2": rel_l2(gx[sig].cpu().numpy(), gx_ref),
                          "max_abs": max_abs(gx[sig].cpu().numpy(), gx_ref)}
    gh_ref = np.zeros(K)
    for c0 in range(0, count, 64):  # chunks keep host memory small
        xc = x[c0:c0 + 64].cpu().numpy().astype(np.float64)
        gc = gy[c0:c0 + 64].cpu().numpy().astype(np.float64)
        gh_ref += fftconvolve(gc, xc[:, ::-1], mode="valid", axes=1).sum(axis=0)
    acc["grad_taps"] = {"rel_l2": rel_l2(gh.cpu().numpy(), gh_ref), "max_abs": max_abs(gh.cpu().numpy(), gh_ref)}
    for name, d in fp16.items():
        o = d.pop("_out")
        acc["forward_" + name] = {"rel_l2": rel_l2(o, y_ref), "max_abs": max_abs(o, y_ref)}
    out = {"signals": count, "samples": S, "taps": K, "fft": N, "accuracy_vs_fp64": acc,
           "forward_ms": fwd, "forward_gbs_alg": fbytes / (fwd * 1e-3) / 1e9,
           "forward_roofline_frac": fbytes / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "backward_ms": bwd, "backward_gbs_alg": bbytes / (bwd * 1e-3) / 1e9,
           "backward_roofline_frac": bbytes / (bwd * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "signals_per_s_fwd_bwd": count / ((fwd + bwd) * 1e-3), "fp16_forward": fp16}
    if cpu:
        # the oracle's FP32 overlap-add restatement (Convolve.swift:36-76),
        # one core, on 16 of the 512 signals
        from oracle import oracle as O
        O.build()
        xs = x[:16].cpu().numpy()
        hh = h.cpu().numpy()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            for r in range(16):
                O.cpu_convolve(xs[r], hh, N)
            ts.append(time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": 16 / float(np.median(ts)), "unit": "signals/s (forward)",
                               "cores": 1, "kind": "port",
                               "sample": "16 of the 512 signals, median of 3"}
    return out


def extra_sar(m, torch, n=4096):
    """Config 5 (range compression -> corner turn -> azimuth compression, no
    RCMC) and the full RDA with RCMC, complex FP16 data, 4096 x 4096; the
    full RDA also with range and azimuth compression in the three FP16
    arithmetic modes."""
    g = torch.Generator(device="cuda").manual_seed(5)
    raw = (torch.randn((n, 2 * n), generator=g, device="cuda") * 0.1).half()
    img = torch.empty_like(raw)
    out = {"pulses": n, "range_samples": n, "io": "complex fp16",
           "accuracy": "checked in tests/test_sar_gpu.py against the FP64 pipeline on the 32-target "
                       "4096 x 4096 scene: rel-L2 1.9e-3 with FP32 arithmetic (the complex-FP16 "
                       "data's own quantization; bound 3e-3), pure / storage / mixed 2.6e-3 / "
                       "2.0e-3 / 2.2e-3 (bounds 6e-3 / 4e-3 / 5e-3)"}
    for key, rcmc, sweeps, prec in (("config5_no_rcmc", 0, 3, 0), ("full_rda_rcmc", 1, 5, 0),
                                    ("full_rda_rcmc_fp16_pure", 1, 5, 1),
                                    ("full_rda_rcmc_fp16_storage", 1, 5, 2),
                                    ("full_rda_rcmc_fp16_mixed", 1, 5, 3)):
        rda = m.RangeDoppler(m.RdaParams.defaults(n_range=n, n_azimuth=n, io=1, rcmc=rcmc,
                                                  precision=prec))
        ms = time_fn(lambda: rda.process(raw, img), 10, 2, torch)
        nbytes = sweeps * 2 * 4 * n * n  # read + write sweeps of complex FP16
        out[key] = {"ms": ms, "frames_per_s": 1e3 / ms, "sweeps": sweeps,
                    "gbs_alg": nbytes / (ms * 1e-3) / 1e9,
                    "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        rda.close()
    return out


def extra_correlate(m, torch, count=4096, n=4096):
    """Correlator.cross (Correlate.swift:27-61) batched on the device: 4096
    pairs of real signals of 4096 samples, N = 4096 circular correlations in
    one launch; accuracy against numpy's FP64 IFFT(A conj(B)) on a sample."""
    g = torch.Generator(device="cuda").manual_seed(12)
    a = torch.randn((count, n), generator=g, device="cuda")
    b = torch.randn((count, n), generator=g, device="cuda")
    r = torch.empty((count, n), device="cuda")
    c = m.Correlator(n)
    ms = time_fn(lambda: c.cross_device(a, n, b, n, count, r), 20, 4, torch)
    c.cross_device(a, n, b, n, count, r)
    torch.cuda.synchronize()
    idx = list(range(0, count, 64))
    a64 = a[idx].cpu().numpy().astype(np.float64)
    b64 = b[idx].cpu().numpy().astype(np.float64)
    ref = np.real(np.fft.ifft(np.fft.fft(a64, axis=1) * np.conj(np.fft.fft(b64, axis=1)), axis=1))
    got = r[idx].cpu().numpy()
    c.close()
    nbytes = 4 * count * (n + n + n)  # a and b read, r written
    return {"pairs": count, "samples": n, "fft": n, "ms": ms, "correlations_per_s": count / (ms * 1e-3),
            "gbs_alg": nbytes / (ms * 1e-3) / 1e9,
            "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "accuracy_vs_fp64": {"rel_l2": rel_l2(got, ref), "max_abs": max_abs(got, ref),
                                 "sample": f"{len(idx)} of the {count} correlations"}}


def extra_xspec(m, torch, c=8, s=1 << 20, n=1024, hop=512):
    """Cross-spectral analysis (PSD.crossSpectral): 8 channels x 2^20 samples,
    Hann window, N = 1024, hop 512 (2047 frames, 28 pairs)."""
    g = torch.Generator(device="cuda").manual_seed(9)
    x = torch.randn((c, s), generator=g, device="cuda")
    w = torch.from_numpy(m.window_coefficients("hann", n)).cuda()
    an = m.CrossSpectral(n, c)
    p = an.pairs
    pw = torch.empty((c, n), device="cuda")
    cr = torch.empty((p, n), dtype=torch.complex64, device="cuda")
    co = torch.empty((p, n), device="cuda")
    ms_ = torch.empty((p, n), device="cuda")
    ms = time_fn(lambda: an.compute_device(x, s, hop, w, pw, cr, co, ms_), 10, 2, torch)
    frames = an.frames(s, hop)
    an.close()
    # algorithmic bytes: frames of every channel read once (overlapping
    # frames re-read), spectra written and read back once
    nbytes = 4 * c * frames * n + 2 * 8 * c * frames * n
    return {"channels": c, "samples": s, "fft": n, "hop": hop, "frames": frames, "pairs": p,
            "ms": ms, "frames_per_s": c * frames / (ms * 1e-3), "gbs_alg": nbytes / (ms * 1e-3) / 1e9,
            "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def extra_rocfft(m, torch, n=4096, batch=4096):
    """Library line (not the reference): torch.fft.fft, i.e. rocFFT through
    hipFFT, on the headline workload (N = 4096, batch 4096, complex64, four
    rotating input sets; torch allocates each output from its cache, and
    out= would add a copy), beside this library on the same sets."""
    sets = [(torch.randn((batch, n), dtype=torch.complex64, device="cuda"),
             torch.empty((batch, n), dtype=torch.complex64, device="cuda")) for _ in range(4)]
    f = m.MetalFFT(n)
    k = [0]

    def rocfft():
        a, _ = sets[k[0] % 4]
        k[0] += 1
        torch.fft.fft(a, dim=1)

    def ours():
        a, b = sets[k[0] % 4]
        k[0] += 1
        f.execute(a, b, batch)
    out = {}
    for name, fn in (("rocfft_torch_fft", rocfft), ("mifft", ours)):
        ms = time_fn(fn, 40, 5, torch)
        out[name] = {"ms": ms, "transforms_per_s": batch / (ms * 1e-3),
                     "roofline_frac": 16 * n * batch / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out["note"] = "stream launches from Python for both (no graph)"
    return out


def extra_sizes(m, torch):
    """Every supported length at 2^24 complex values per launch (128 MiB in,
    128 MiB out), four rotating sets, the autotuned kernel, graph replay of
    200 launches: transforms/s and fraction of the HBM peak (algorithmic
    bytes, 16 n per transform; the four-step lengths 32768 ... 2^20 move
    twice that through HBM in their two passes)."""
    import ctypes
    from metalfft_amd import _lib
    L = _lib.load()
    out = {}
    for n in m.SUPPORTED_SIZES + m.FOURSTEP_SIZES:
        batch = (1 << 24) // n
        f = m.MetalFFT(n)
        sets = [(torch.randn((batch, n), dtype=torch.complex64, device="cuda"),
                 torch.empty((batch, n), dtype=torch.complex64, device="cuda")) for _ in range(4)]

        def step(i, on_stream, f=f, sets=sets, batch=batch):
            src, dst = sets[i % 4]
            st = ctypes.c_void_p(on_stream if on_stream is not None
                                 else torch.cuda.current_stream().cuda_stream)
            _lib.check(L.mifft_execute_device(f._plan, ctypes.c_void_p(src.data_ptr()),
                                              ctypes.c_void_p(dst.data_ptr()), ctypes.c_int64(batch),
                                              ctypes.c_int(m.FORWARD), st))
        ms, _, graph, _, _ = time_graph(step, 200, 8, torch)
        del graph
        gbs = 16 * n * batch / (ms * 1e-3) / 1e9
        out[str(n)] = {"kernel": f.kernel_name, "us": ms * 1e3, "transforms_per_s": batch / (ms * 1e-3),
                       "gbs": gbs, "roofline_frac": gbs / HBM_PEAK_GBS}
        del sets
        torch.cuda.empty_cache()
    return out


def extra_real(m, torch, n=4096, batch=4096):
    """Real-input transforms, 4096 real signals of 4096 samples: forward to
    all N bins, to bins 0..N/2, and the inverse from bins 0..N/2."""
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn((batch, n), generator=g, device="cuda")
    full = torch.empty((batch, n), dtype=torch.complex64, device="cuda")
    half = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda")
    back = torch.empty_like(x)
    f = m.MetalFFT(n)
    out = {"fft": n, "batch": batch}
    for key, fn, nbytes in (
            ("forward_full", lambda: f.forward_real(x, full, batch), (4 * n + 8 * n) * batch),
            ("forward_half", lambda: f.forward_real(x, half, batch, half=True),
             (4 * n + 8 * (n // 2 + 1)) * batch),
            ("inverse_half", lambda: f.inverse_real(half, back, batch, half=True),
             (4 * n + 8 * (n // 2 + 1)) * batch)):
        ms = time_fn(fn, 20, 3, torch)
        out[key] = {"ms": ms, "transforms_per_s": batch / (ms * 1e-3),
                    "gbs_alg": nbytes / (ms * 1e-3) / 1e9,
                    "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return out


def extra_stft(m, torch, s=1 << 24, n=1024, hop=256):
    """Batched framing: STFT.analyze of one 2^24-sample signal, Hann window,
    N = 1024, hop 256 (65533 frames in one launch; spectra and |X|^2
    written), and PSD.welch of the same signal."""
    g = torch.Generator(device="cuda").manual_seed(10)
    x = torch.randn((1, s), generator=g, device="cuda")
    w = torch.from_numpy(m.window_coefficients("hann", n)).cuda()
    an = m.CrossSpectral(n, 1)
    f = an.frames(s, hop)
    spec = torch.empty((f, n), dtype=torch.complex64, device="cuda")
    pw = torch.empty((f, n), device="cuda")
    psd = torch.empty(n, device="cuda")
    ms = time_fn(lambda: an.stft_device(x, s, hop, w, spec, pw), 10, 2, torch)
    ms_w = time_fn(lambda: an.welch_device(x, s, hop, w, 44100.0, psd), 10, 2, torch)
    an.close()
    # algorithmic bytes: the signal read once, spectra (8 B) and |X|^2 (4 B)
    # per bin written
    nbytes = 4 * s + 12 * f * n
    return {"samples": s, "fft": n, "hop": hop, "frames": f, "stft_ms": ms,
            "frames_per_s": f / (ms * 1e-3), "gbs_alg": nbytes / (ms * 1e-3) / 1e9,
            "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "welch_ms": ms_w,
            "welch_frames_per_s": f / (ms_w * 1e-3)}


def load_traffic(name, kernel):
    """HBM bytes per launch from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary, only if it was measured on the kernel variant now timed;
    (bytes or None, note)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            d = json.load(f)
    except Exception:
        return None, f"no traffic summary {name}"
    measured = d.get("variant") or d.get("kernel", "").split(" ")[0]
    if measured != kernel:
        return None, f"traffic in {name} was measured on {measured}, the timed kernel is {kernel}"
    return d.get("hbm_bytes_per_launch"), f"{name} (measured on {measured})"


# --probe workloads: a few launches each, nothing timed (rocprofv3 --pmc runs)
PROBES = {
    "fft4096": "batched FFT N=4096, batch 4096, complex64 (the headline kernel)",
    "fft8192": "batched FFT N=8192, batch 2048, complex64",
    "fft16384": "batched FFT N=16384, batch 1024, complex64",
    "mfma_fp16split_c64": "DFT-as-GEMM FP16 hi/lo, N=4096, batch 4096, complex64 I/O",
    "mfma_fp16split_c32h": "DFT-as-GEMM FP16 hi/lo, N=4096, batch 4096, complex-FP16 I/O",
    "sar_fp16": "range-Doppler 4096 x 4096 complex FP16, RCMC on, FP32 arithmetic",
    "sar_fp16_config5": "range-Doppler 4096 x 4096 complex FP16, no RCMC (config 5), FP32 arithmetic",
    "convolve": "fused convolution 512 x 65536, K = 1023, N = 4096, forward + adjoint",
    "convolve_fp16": "fused convolution 512 x 65536, K = 1023, N = 4096, forward in the FP16 modes "
                     "pure, storage, mixed",
}


def run_probe(name, torch, m, launches=4):
    g = torch.Generator(device="cuda").manual_seed(0)
    done = []
    if name.startswith("fft"):
        n = int(name[3:])
        batch = (1 << 24) // n
        f = m.MetalFFT(n)
        x = torch.randn((batch, n), dtype=torch.complex64, device="cuda", generator=g)
        y = torch.empty_like(x)
        for _ in range(launches):
            f.execute(x, y, batch)
        done.append(f.kernel_name)
    elif name.startswith("mfma"):
        io = 1 if name.endswith("c32h") else 0
        p = m.PrecisionFFT(4096, m.PrecisionFFT.FP16_MFMA_SPLIT, io)
        x = torch.randn((4096, 4096), dtype=torch.complex64, device="cuda", generator=g)
        if io:
            x = torch.view_as_real(x).reshape(4096, 8192).half().contiguous()
        y = torch.empty_like(x)
        for _ in range(launches):
            p.execute(x, y, 4096)
        done.append(p.kernel_name)
    elif name.startswith("sar_fp16"):
        rcmc = 0 if name.endswith("config5") else 1
        rda = m.RangeDoppler(m.RdaParams.defaults(n_range=4096, n_azimuth=4096, io=1, rcmc=rcmc))
        raw = (torch.randn((4096, 8192), generator=g, device="cuda") * 0.1).half()
        img = torch.empty_like(raw)
        for _ in range(launches):
            rda.process(raw, img)
        done.append("k_rows_stream / k_corner_turn_h16" + (" / k_rcmc" if rcmc else ""))
    elif name == "convolve":
        K, S, cnt = 1023, 65536, 512
        h = torch.randn(K, generator=g, device="cuda") / np.sqrt(K)
        cv = m.FusedConvolver(h.cpu().numpy(), 4096)
        x = torch.randn((cnt, S), generator=g, device="cuda")
        y = torch.empty((cnt, S + K - 1), device="cuda")
        gy = torch.randn_like(y)
        gx = torch.empty_like(x)
        gh = torch.empty(K, device="cuda")
        for _ in range(launches):
            cv.apply_device(x, y, S, cnt)
            cv.backward_device(x, gy, gx, gh, S, cnt)
        done.append("k_conv_fwd / k_conv_bwd_x / k_conv_bwd_h / k_sum_rows64")
    elif name == "convolve_fp16":
        K, S, cnt = 1023, 65536, 512
        h = (torch.randn(K, generator=g, device="cuda") / np.sqrt(K)).cpu().numpy()
        x = torch.randn((cnt, S), generator=g, device="cuda")
        y = torch.empty((cnt, S + K - 1), device="cuda")
        for prec in ("float16Pure", "float16Storage", "float16Mixed"):
            cv = m.FusedConvolver(h, 4096, getattr(m.Precision, prec))
            for _ in range(launches):
                cv.apply_device(x, y, S, cnt)
        done.append("k_conv_fwd_h<1> / <2> / <3>")
    torch.cuda.synchronize()
    print(json.dumps({"probe": name, "workload": PROBES[name], "launches": launches, "kernels": done}))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(transforms are independent: N GPUs would each run this batch)", file=sys.stderr)
        sys.exit(2)
    import torch
    import metalfft_amd as m

    if a.probe:
        run_probe(a.probe, torch, m)
        return

    n, batch = a.n, a.batch
    fft = m.MetalFFT(n)
    if a.kernel:
        fft.select_kernel(a.kernel)
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    sets = []
    for s in range(a.sets):
        x = (torch.rand((batch, n, 2), generator=g, device="cuda") * 2 - 1)
        sets.append((torch.view_as_complex(x.contiguous()), torch.empty((batch, n), dtype=torch.complex64,
                                                                      device="cuda")))
    torch.cuda.synchronize()

    # launch through a pre-bound ctypes call: per-step Python work is one
    # foreign call, so back-to-back launches keep the GPU fed
    import ctypes
    from metalfft_amd import _lib
    L = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    calls = [(fft._plan, ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(d.data_ptr()),
              ctypes.c_int64(batch), ctypes.c_int(m.FORWARD), stream) for s, d in sets]
    execute = L.mifft_execute_device

    def step(i, on_stream):
        c = calls[i % len(calls)]
        if on_stream is not None:
            c = c[:5] + (ctypes.c_void_p(on_stream),)
        rc = execute(*c)
        if rc != 0:
            _lib.check(rc)

    if a.traffic_probe:
        # copy reference: complex64 out = in * 1 over the same 128 MiB sets,
        # 128 MiB read + 128 MiB written per launch
        for i in range(4):
            src, dst = sets[i % len(sets)]
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        for i in range(4):
            step(i, None)
        torch.cuda.synchronize()
        print(json.dumps({"traffic_probe": True, "kernel": fft.kernel_name,
                          "copy_bytes_read": 8 * n * batch, "copy_bytes_written": 8 * n * batch}))
        return

    # the reported form first (each form has its own warm-up): the other
    # form's 200 launches do not run right before the timed region
    if a.launch == "graph":
        ms_graph, wall_graph, graph, replays, settled = time_graph(
            step, a.steps, a.warmup, torch, reps=a.replays, settle_s=a.settle)
        del graph
        ms_loop, wall_loop = time_gpu(step, a.steps, a.warmup, torch)
    else:
        ms_loop, wall_loop = time_gpu(step, a.steps, a.warmup, torch)
        ms_graph, wall_graph, graph, replays, settled = time_graph(
            step, a.steps, a.warmup, torch, reps=a.replays, settle_s=a.settle)
        del graph
    ms, wall = (ms_graph, wall_graph) if a.launch == "graph" else (ms_loop, wall_loop)
    copy = copy_ceiling(sets, a.steps, torch)
    value = batch / (ms * 1e-3)
    alg_bytes = 16 * n * batch
    achieved = alg_bytes / (ms * 1e-3) / 1e9

    # accuracy of the measured configuration on a sample (vs numpy's FP64 FFT)
    src, dst = sets[0]
    fft.execute(src, dst, batch=batch)
    torch.cuda.synchronize()
    idx = torch.arange(0, batch, 97, device="cuda")
    xs = src[idx].cpu().numpy()
    ref64 = np.fft.fft(xs.astype(np.complex128), axis=-1)
    err = rel_l2(dst[idx].cpu().numpy(), ref64)
    err_max = max_abs(dst[idx].cpu().numpy(), ref64)

    cpu = None
    if not a.no_cpu:
        xh = src.cpu().numpy()
        cpu = cpu_baseline(n, batch, xh)

    extras = {}
    if not a.no_extras:
        for key, fn in (("sizes_128MiB", lambda: extra_sizes(m, torch)),
                        ("rocfft_4096x4096", lambda: extra_rocfft(m, torch)),
                        ("mfma", lambda: extra_mfma(m, torch)),
                        ("ragged_10k", lambda: extra_ragged(m, torch, cpu=not a.no_cpu)),
                        ("convolve_512x65536_k1023",
                         lambda: extra_convolve(m, torch, cpu=not a.no_cpu)),
                        ("sar_4096x4096_fp16", lambda: extra_sar(m, torch)),
                        ("correlate_4096x4096", lambda: extra_correlate(m, torch)),
                        ("cross_spectral_8ch", lambda: extra_xspec(m, torch)),
                        ("stft_welch_2p24", lambda: extra_stft(m, torch)),
                        ("real_fft_4096x4096", lambda: extra_real(m, torch))):
            if a.only and key not in a.only.split(","):
                continue
            try:
                extras[key] = fn()
            except Exception as e:  # keep the headline line whatever happens here
                extras[key] = {"error": repr(e)}
            torch.cuda.synchronize()

    traffic, traffic_note = load_traffic(f"traffic_stockham_n{n}_b{batch}.json", fft.kernel_name)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: complex64 re/im i.i.d. U[-1,1) (torch generator seed 0), "
                f"{a.sets} rotating buffer sets ({a.sets * 2 * alg_bytes // 2**20} MiB)",
        "config": {"workload": f"batched complex FFT forward, N={n}, batch {batch}, complex64 "
                               "interleaved, FP32 Stockham in LDS",
                   "n": n, "batch": batch, "kernel": fft.kernel_name},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "copy_gbs": copy["gbs"],
                     "frac_of_copy": achieved / copy["gbs"],
                     "traffic": traffic, "traffic_source": traffic_note},
        "cpu_baseline": cpu,
        "rel_l2_vs_fp64": err,
        "max_abs_vs_fp64": err_max,
        "gflops_5nlog2n": 5 * n * np.log2(n) * value / 1e9,
        "device": m.device_name(),
        "wall_s_timed": wall,
        "launch": {"timed": (f"hipGraph of the K launches: {settled} untimed replays until the "
                             f"replay time stopped falling, then {a.replays} replays back to back, "
                             "each between its own HIP events on the replay stream; the median")
                   if a.launch == "graph" else "K launches from a Python loop",
                   "graph_replay_ms_per_step": ms_graph, "replays_ms_per_step": replays,
                   "replay_min_max_ms_per_step": [min(replays), max(replays)],
                   "settle_replays": settled, "stream_loop_ms_per_step": ms_loop},
        "extras": extras,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
