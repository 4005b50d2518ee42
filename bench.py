"""Benchmark: batched complex FFT N=4096, batch 4096, complex64, on one GPU.

One step = one batched forward transform of 4096 length-4096 complex64
signals (128 MiB in, 128 MiB out) already resident in HBM.  Input sets
rotate over >= 1 GiB so the 256 MiB Infinity Cache cannot hold a step's
working set.  Prints one JSON line (driver contract); extra keys carry the
roofline, the CPU baseline and the reduced-precision forms.

    python bench.py [--gpus 1] [--steps 50] [--warmup 10]
"""
from __future__ import annotations

import argparse
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
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--n", type=int, default=4096)
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--sets", type=int, default=4, help="rotating in/out buffer sets")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--kernel", default=None, help="pin a kernel variant by name")
    return p.parse_args()


def time_gpu(fn, steps, warmup, torch):
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(steps):
        fn(i)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


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


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(transforms are independent: N GPUs would each run this batch)", file=sys.stderr)
        sys.exit(2)
    import torch
    import metalfft_amd as m

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

    def step(i):
        src, dst = sets[i % len(sets)]
        fft.execute(src, dst, batch=batch)

    ms, wall = time_gpu(step, a.steps, a.warmup, torch)
    value = batch / (ms * 1e-3)
    alg_bytes = 16 * n * batch
    achieved = alg_bytes / (ms * 1e-3) / 1e9

    # accuracy of the measured configuration on a sample (vs FP64)
    from oracle import oracle as O
    src, dst = sets[0]
    fft.execute(src, dst, batch=batch)
    torch.cuda.synchronize()
    idx = torch.arange(0, batch, 97, device="cuda")
    xs = src[idx].cpu().numpy()
    err = O.rel_l2(dst[idx].cpu().numpy(), O.f64_fft(xs))

    cpu = None
    if not a.no_cpu:
        xh = src.cpu().numpy()
        cpu = cpu_baseline(n, batch, xh)

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
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(f"traffic_stockham_n{n}_b{batch}.json")},
        "cpu_baseline": cpu,
        "rel_l2_vs_fp64": err,
        "gflops_5nlog2n": 5 * n * np.log2(n) * value / 1e9,
        "device": m.device_name(),
        "wall_s_timed": wall,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
