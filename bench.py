"""Benchmark: particle-steps/s on one MI355X, batched vs plain, with step
latency percentiles, the roofline of the dominant kernel and a CPU baseline.

Workload: 256 independent dam-break / colliding-drop
scenes with particle counts log-uniform in [1e3, 2e5] (default_rng(3)),
delta-SPH (RK4, mDBC walls) and DFSPH (semi-implicit, Akinci walls) 50/50,
fp32 state, one hipGraph-captured step per time step per scheme.
A step advances every scene by one time step; value = sum of particles /
step time.  "plain" is the one-scene-at-a-time path: the same kernels on a
batch of one scene, timed on a random sample of the scenes.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="run the delta-SPH and DFSPH parts one after the other (profiling)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_roofline(batch):
    """Average duration of the fused delta-SPH gather (the dominant kernel)
    over repeated launches on the stream it runs on, and its algorithmic
    HBM bytes: per fluid particle, own state x, v, rho, A (24 B) + visit
    order and scene id (8 B) + row pointers of the fluid and boundary lists
    (8 B) + its K column indices (4 K B) + dpdt and acc written (12 B);
    neighbour state re-reads are cache traffic, not compulsory bytes."""
    import torch

    from pytorchsph_amd import _lib as L

    b = batch
    b.dsph_rhs(b.x, b.v, b.rho)          # pair lists of the current state
    torch.cuda.synchronize()
    nnz = int(b.ff.rowptr[b.NP].item())
    nnz_b = int(b.fb.rowptr[b.NP].item())
    F = b.fluid(b.x, b.v, b.rho)
    ff, fb, B = b.ff.csr().c(), b.fb.csr().c(), b.boundary()
    stream = torch.cuda.current_stream()
    sh = L.stream_handle(stream)

    def launch():
        L.call("sph_dsph_gather", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B),
               C.byref(fb), None, b.terms, L.ptr(b.grid.sorted_idx), L.ptr(b.dpdt),
               L.ptr(b.acc), sh)

    for _ in range(3):
        launch()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    e1.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    n = b.NP
    bytes_alg = n * (24 + 8 + 8 + 12) + 4 * nnz + 4 * nnz_b
    achieved = bytes_alg / sec / 1e9
    return dict(kernel="k_dsph_gather (sph_dsph_gather)", seconds=sec, bytes=bytes_alg,
                particles=n, pairs=nnz, achieved_gbs=achieved)


def cpu_baseline(seconds: float):
    """The oracle's delta-SPH RK4 step (fp32 neighbour search, FP64 physics,
    OpenMP) on the 5,000-particle dam break, on this host's cores."""
    from oracle import oracle as O
    from pytorchsph_amd import workloads as W

    threads = os.cpu_count() or 1
    O.set_threads(threads)
    sc = W.c1_scene()
    so = O.SceneOracle(sc)
    x, v, r = sc.x.astype(np.float64), sc.v.astype(np.float64), np.ones(sc.n)
    for _ in range(3):
        x, v, r = so.dsph_step(x, v, r)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end and len(times) < 1000:
        t0 = time.perf_counter()
        x, v, r = so.dsph_step(x, v, r)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return dict(value=sc.n / med, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"oracle delta-SPH RK4 dam break, {sc.n} fluid + {sc.nb} boundary "
                       f"particles, median of {len(times)} steps after 3 warm-up")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("multi-GPU runs are not part of this build: scenes shard across GPUs with no "
            "exchange (see DESIGN.md); run with --gpus 1")
        sys.exit(2)
    import torch

    from pytorchsph_amd import workloads as W
    from pytorchsph_amd.engine import MixedBatch, SceneBatch

    t0 = time.perf_counter()
    scenes = W.c4_scenes(n_scenes=args.scenes)
    n_part = int(sum(s.n for s in scenes))
    log(f"{len(scenes)} scenes, {n_part} fluid particles, generated in "
        f"{time.perf_counter() - t0:.1f}s")
    batch = MixedBatch(scenes)
    batch.capture()
    torch.cuda.synchronize()
    log("captured")
    for _ in range(args.warmup):
        batch.step(serial=args.serial)
    torch.cuda.synchronize()
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        batch.step(serial=args.serial)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        log(f"step {k}: {lat[-1] * 1e3:.1f} ms")
    total = time.perf_counter() - t_start
    ms_per_step = total / args.steps * 1e3
    value = n_part * args.steps / total
    status = batch.status()
    iters = {b.scheme: b.dfsph_iterations().tolist() for b in batch.parts if b.scheme == "dfsph"}

    rl = kernel_roofline(next(b for b in batch.parts if b.scheme == "dsph"))
    log(f"roofline: {rl}")
    del batch
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain:
        rng = np.random.default_rng(7)
        pick = rng.choice(len(scenes), size=min(args.plain_sample, len(scenes)), replace=False)
        t_sum, n_sum = 0.0, 0
        for k in pick:
            b = SceneBatch([scenes[k]])
            b.capture()
            b.step_graph()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 2
            for _ in range(reps):
                b.step_graph()
            torch.cuda.synchronize()
            t_sum += (time.perf_counter() - t0) / reps
            n_sum += scenes[k].n
            del b
        plain = n_sum / t_sum
        log(f"plain one-scene-at-a-time: {plain:.3e} particle-steps/s over {len(pick)} scenes")

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    peak = 8000.0
    out = {
        "metric": "particle-steps/s on 1 GPU, batched vs plain; p50/p99 step latency under load",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: 256 dam-break/colliding-drop scenes, 1e3-2e5 particles "
                               "log-uniform (seed 3), delta-SPH RK4 / DFSPH 50/50, "
                               "hipGraph-captured step",
                   "scenes": len(scenes), "particles": n_part, "parallelism": "single GPU"},
        "plain_value": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "p50_step_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_step_ms": float(np.percentile(lat, 99) * 1e3),
        "dfsph_iterations_mean": float(np.mean(iters["dfsph"])) if iters else None,
        "status_flags": int(np.bitwise_or.reduce(status)),
        "roofline": {"bound": "hbm", "achieved": rl["achieved_gbs"], "peak": peak,
                     "unit": "GB/s", "frac": rl["achieved_gbs"] / peak, "traffic": None,
                     "kernel": rl["kernel"], "avg_launch_us": rl["seconds"] * 1e6,
                     "bytes_per_launch": rl["bytes"], "particles": rl["particles"],
                     "pairs": rl["pairs"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
