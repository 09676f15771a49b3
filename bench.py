"""Benchmark: particle-steps/s on one MI355X, batched vs plain, with step
latency percentiles, the roofline of the dominant kernel and a CPU baseline.

Workload (the batched benchmark): 256 independent dam-break / colliding-drop
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=200,
                    help="untimed steps before the timed window (the scenes flow first)")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="run the delta-SPH and DFSPH parts one after the other (profiling)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-paths", action="store_true",
                    help="skip the single-scene workloads (C2 search, 4M delta-SPH, C3 DFSPH)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the serving run (0 = skip)")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shapes instead of autotuning: count,gather,solve[,fill"
                         "[,walk[,block]]] lanes per row (block: workgroup size) and the DFSPH "
                         "grid cap, e.g. 2,4,2,8,1,256,16384")
    ap.add_argument("--path", choices=PATHS, default=None,
                    help="run one single-scene workload of the plain paths and print it")
    ap.add_argument("--dplus-only", action="store_true",
                    help="only the delta-plus surface detection and shift timings (profiling)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the kernel timings and the calibration kernels (counter runs)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _events_avg(launch, stream, reps=20, warm=3):
    """Average duration of `launch` over `reps` launches, timed with events
    recorded on the stream the kernel runs on."""
    import torch

    for _ in range(warm):
        launch()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def pressure_roofline(b, reps=20):
    """The DFSPH pressure update with every DFSPH scene active (after one
    accel, whose records it reads): sph_dfsph_pressure_sl (k_pressure_sl)
    on a stencil-list batch, else sph_dfsph_pressure (k_pressure), and its
    algorithmic HBM bytes:
      per fluid row      76 B: fluid and boundary row pointers (4 + 4), own
                         xrec and arec records (16 + 16), src, p2, alpha,
                         rho, rest (20), p, p2, resid written (12), xrec.z
                         written (4); with stencil lists no fluid row
                         pointer but the row's runs and masks (40)
      per boundary row   56 B: row pointer (4), bxrec (16), b_aact (4),
                         b_src, b_p, b_alpha, rho_b, rest_b (20), b_p,
                         b_resid written (8), bxrec.z written (4)
      per pair           4 B column index (fluid-fluid with a CSR list,
                         fluid-boundary, boundary-fluid lists)
    Neighbour records are counted once (as their owner's row), so re-reads
    of a neighbour are cache traffic, not algorithmic bytes."""
    import ctypes as C

    import torch

    from pytorchsph_amd import _lib as L

    b._dfsph_pre()              # pair lists, density, alpha/source; all scenes active
    torch.cuda.synchronize()
    sl = b.stencil_list() if b.stencil_lists else None
    nnz_ff = 0 if sl is not None else int(b.ff.rowptr[b.NP].item())
    nnz_fb = int(b.fb.rowptr[b.NP].item()) if b.NB else 0
    nnz_bf = int(b.bf.rowptr[b.NB].item()) if b.NB else 0
    F = b.fluid(b.x, b.v, b.rho)
    ff = b.ff.csr().c()
    fb = b.fb.csr().c() if b.NB else None
    bf = b.bf.csr().c()
    B = b.boundary()
    D = b._dfsph_struct()
    stream = torch.cuda.current_stream()
    sh = L.stream_handle(stream)

    fbp = C.byref(fb) if fb is not None else None
    if sl is not None:
        L.call("sph_dfsph_accel_sl", C.byref(F), L.ptr(b.phys), C.byref(sl), fbp, C.byref(D), sh)
    else:
        L.call("sph_dfsph_accel", C.byref(F), L.ptr(b.phys), C.byref(ff), fbp, C.byref(D), sh)

    def launch():
        if sl is not None:
            L.call("sph_dfsph_pressure_sl", C.byref(F), L.ptr(b.phys), C.byref(sl), C.byref(B),
                   fbp, C.byref(bf), C.byref(D), sh)
        else:
            L.call("sph_dfsph_pressure", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B),
                   fbp, C.byref(bf), C.byref(D), sh)

    sec = _events_avg(launch, stream, reps)
    row = 76 + (36 if sl is not None else 0)
    nbytes = row * b.NP + 56 * b.NB + 4 * (nnz_ff + nnz_fb + nnz_bf)
    name = ("k_pressure_sl (sph_dfsph_pressure_sl)" if sl is not None
            else "k_pressure (sph_dfsph_pressure)")
    return dict(kernel=name, seconds=sec, bytes=nbytes,
                rows=b.NP + b.NB, pairs=nnz_ff + nnz_fb + nnz_bf,
                achieved_gbs=nbytes / sec / 1e9)


def gather_roofline(b, reps=20):
    """sph_dsph_gather_packed (k_dsph_gather_packed, all delta-SPH terms)
    over the fluid pair list, the gather the cell walk replaces, and its
    algorithmic HBM bytes (each array once): per fluid row x, v, rho, A (24),
    dpdt and a written (12) and its fluid and boundary row pointers (8); per
    pair its column (4); per boundary point x, rho_b, V_b, p_b (20).  Its
    packed records (32 B per fluid row, 32 per boundary point) and the scene
    ids (4 per row) are reported apart as `record_bytes`."""
    import ctypes as C

    import torch

    from pytorchsph_amd import _lib as L

    b.search(b.x)                    # every pair list of the current state
    b.dsph_rhs(b.x, b.v, b.rho)      # mDBC and records of the current state
    torch.cuda.synchronize()
    nnz = int(b.ff.rowptr[b.NP].item())
    nnz_b = int(b.fb.rowptr[b.NP].item())
    ff, fb = b.ff.csr().c(), b.fb.csr().c()
    stream = torch.cuda.current_stream()
    sh = L.stream_handle(stream)

    def launch():
        L.call("sph_dsph_gather_packed", L.ptr(b.frec), L.ptr(b.brec), L.ptr(b.pscene), b.NP,
               L.ptr(b.phys), C.byref(ff), C.byref(fb),
               L.ptr(b.gradrho) if b.terms & L.T_DDSPH else None, b.terms, b._order(),
               L.ptr(b.dpdt), L.ptr(b.acc), sh)

    sec = _events_avg(launch, stream, reps)
    nbytes = 44 * b.NP + 20 * b.NB + 4 * (nnz + nnz_b)
    return dict(kernel="k_dsph_gather_packed (sph_dsph_gather_packed)", seconds=sec,
                bytes=nbytes, record_bytes=36 * b.NP + 32 * b.NB, rows=b.NP,
                pairs=nnz + nnz_b, achieved_gbs=nbytes / sec / 1e9)


def cellwalk_roofline(b, reps=20):
    """sph_dsph_cellwalk_rk4 (k_dsph_cellwalk with the MOG delta-SPH terms
    and the RK4 stage fused), the dominant kernel of the delta-SPH step:
    stage 1 of the current state, with its algorithmic HBM bytes (each array
    once):
      per fluid row      x, v, rho, A (24) + the stencil's sorted positions
                         and point indices (12) + the RK4 stage combine (52:
                         x, v, rho of the step start 20, running sums 12,
                         new state 20)
      per cell           its start in the fluid and in the boundary cell
                         table (4 + 4)
      per boundary point its record (32) + its sorted position and index
                         in the static boundary grid (12)
    The kernel reads each fluid neighbour through a 32-byte packed record
    {x, y, vx, vy, rho, V, p, rho_hat}; those bytes (32 per fluid row) are
    reported apart as `record_bytes`.  It reads no pair list."""
    import ctypes as C

    import torch

    from pytorchsph_amd import _lib as L

    assert b.cellwalk and b.fuse_rk4, "a cell-walk batch with the fused stage"
    b._dsph_prepare(b.x, b.v, b.rho, b.frec)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    sh = L.stream_handle(stream)
    s0 = b.S[0]
    rk = L.SphRk4Stage(1, L.ptr(b.x), L.ptr(b.v), L.ptr(b.rho), L.ptr(b.area), L.ptr(b.K),
                       L.ptr(s0[0]), L.ptr(s0[1]), L.ptr(s0[2]), None, L.ptr(b.frec2))
    g = b.grid._grid()
    bg = b._bgrid_arg()

    def launch():
        L.call("sph_dsph_cellwalk_rk4", C.byref(g), L.ptr(b.segs), b.nseg, L.ptr(b.frec),
               L.ptr(b.brec), L.ptr(b.phys), None, bg, b.terms, C.byref(rk), sh)

    sec = _events_avg(launch, stream, reps)
    bcells = b.bgrid.ncells if b.bgrid is not None else 0
    nbytes = (24 + 12 + 52) * b.NP + 4 * (b.grid.ncells + bcells) + 44 * b.NB
    # the boundary pairs the launch found, for the record (one search)
    b.search(b.x)
    torch.cuda.synchronize()
    nnz_b = int(b.fb.rowptr[b.NP].item()) if b.NB else 0
    return dict(kernel="k_dsph_cellwalk (sph_dsph_cellwalk_rk4)", seconds=sec, bytes=nbytes,
                record_bytes=32 * b.NP, rows=b.NP, boundary_pairs=nnz_b,
                achieved_gbs=nbytes / sec / 1e9)


def read_write_calibration(n_ints=1 << 26):
    """Two more kernels of known traffic for the counter scale factors: a
    read-only pass (k_scan_partials of sph_scan_exclusive reads 256 MiB of
    counts and writes 128 KiB of partial sums) and a write-only pass
    (torch's fill of 256 MiB)."""
    import torch

    from pytorchsph_amd import _lib as L
    from pytorchsph_amd.search import scan

    cnt = torch.ones(n_ints, dtype=torch.int32, device="cuda")
    out = torch.empty(n_ints + 1, dtype=torch.int32, device="cuda")
    for _ in range(3):
        scan(cnt, n_ints, out)
    fill = torch.empty(n_ints, dtype=torch.float32, device="cuda")
    for _ in range(3):
        fill.fill_(2.0)
    torch.cuda.synchronize()
    return dict(read_only_bytes=4 * n_ints, write_only_bytes=4 * n_ints)


def copy_calibration(n_floats=1 << 26):
    """A plain copy of known size through the library's copy kernel
    (sph_permute_f32 without a permutation): 256 MiB read, 256 MiB written.
    Under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    import torch

    from pytorchsph_amd import _lib as L

    src = torch.ones(n_floats, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        L.call("sph_permute_f32", L.ptr(src), None, n_floats, 1, L.ptr(dst), L.stream_handle())
    torch.cuda.synchronize()
    return 4 * n_floats


TRAFFIC_PROFILE = "profiles/roofline_kernels_traffic_c4.md"
ISSUE_PROFILE = "profiles/roofline_kernels_counters_c4.md"


def _profile_value(rel: str, key: str, kernel: str):
    """`key=<number> kernel=<name>` from a committed counter summary (the
    rocprofv3 runs of `bench.py --roofline-only` it records); None when the
    file or the line is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), rel)
    if not os.path.exists(path):
        return None
    want = kernel.split()[0]
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + "=") and f"kernel={want}" in line:
                return float(line.split("=")[1].split()[0])
    return None


def traffic_from_profile(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE of rocprofv3
    --pmc, scaled by the calibration kernels), from the committed summary of
    that measurement: counters cannot be read from inside this process."""
    v = _profile_value(TRAFFIC_PROFILE, "traffic_bytes_per_launch", kernel)
    return int(v) if v is not None else None


def valu_issue_from_profile(kernel: str):
    """Fraction of `kernel`'s cycles its VALU instructions need at one wave
    instruction per SIMD every 4 cycles (SQ_INSTS_VALU against
    GRBM_GUI_ACTIVE), from the committed counter summary."""
    return _profile_value(ISSUE_PROFILE, "valu_issue_frac", kernel)


def _timed_steps(b, warmup: int, steps: int):
    import torch
    b.capture()
    for _ in range(warmup):
        b.step_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step_graph()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


PATHS = ("search_c2", "dsph_4m", "dfsph_c3", "dsph_c1", "dplus_4m")


def plain_paths(steps: int = 3, warmup: int = 1, which=None):
    """The single-scene workloads, each on its own:
      search_c2     neighbour search of 4,000,000 points (half uniform, half
                    8 Gaussian clusters, ~20 mean neighbours, default_rng(1)):
                    grid build + count + scan + fill, timed with events
      dsph_4m       plain delta-SPH RK4 step of a 2000 x 2000 dam break
                    (dx = 0.0005, 4,000,000 fluid particles, mDBC walls)
      dfsph_c3      plain DFSPH step of the 1,000,000-particle dam break
                    (dx = 0.001, Akinci walls), solved to tau = 1e-3, and
                    the pressure update's roofline on it (every row active)
      dsph_c1       the 5,000-particle delta-SPH dam break of the CPU
                    baseline, one graph-captured step per synchronise
                    (latency-bound at this size)
      dplus_4m      delta-plus surface detection, shift and pressure switch
                    on a 4,000,000-particle dam break (dplus_path)
    `which` names the ones to run (default: all of PATHS).
    Algorithmic bytes (compulsory, each array once; K = mean neighbours,
    c = cells per point):
      search        x, h read (12) + sorted x, index written (12) + key (4)
                    + cell tables (8 c) + rowptr (4) + columns (4 K)
      delta-SPH     per right-hand side: search + the gather's row pointers
                    (4) + columns (4 K) + x, v, rho, A (24) + dpdt, a (12);
                    per RK stage combine 52; four of each per step: the
                    compulsory bytes of the pair-list algorithm with one fused
                    gather per right-hand side (the engine's cell walk moves
                    fewer: it writes and reads no fluid list)"""
    import torch

    from pytorchsph_amd import _lib as L
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd import workloads as W
    from pytorchsph_amd.engine import SceneBatch
    from pytorchsph_amd.search import NeighborGrid, grid_params, scan

    which = set(which or PATHS)
    out = {}
    dev = torch.device("cuda")
    peak = 8000.0
    if "search_c2" in which:
        # ---- C2: search alone
        x, h = W.c2_points()
        n = len(x)
        g = NeighborGrid([dict(vmin=(0.0, 0.0), vmax=(1.0, 1.0), hmax=np.float32(h), p_begin=0,
                               p_count=n)], n, dev)
        xt = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
        ht = torch.full((n,), float(h), dtype=torch.float32, device=dev)
        ps = torch.zeros(n, dtype=torch.int32, device=dev)
        counts = torch.empty(n, dtype=torch.int32, device=dev)
        rowptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        g.build(xt, ps, ordered=False)
        q = g.query_struct(xt, ht, ps, 1, True)
        g.count(q, counts)
        scan(counts, n, rowptr)
        nnz = int(rowptr[n].item())
        col = torch.empty(nnz, dtype=torch.int32, device=dev)

        def one_search():
            g.build(xt, ps, ordered=False)
            g.count(q, counts)
            scan(counts, n, rowptr)
            g.fill(q, rowptr, col)

        sec = _events_avg(one_search, torch.cuda.current_stream(), reps=10, warm=2)
        cells = g.ncells / n
        k_mean = nnz / n
        nbytes = n * (12 + 12 + 4 + 8 * cells + 4 + 4 * k_mean)
        out["search_c2"] = dict(points=n, pairs=nnz, ms=sec * 1e3, searches_per_s=1 / sec,
                                achieved_gbs=nbytes / sec / 1e9, frac=nbytes / sec / 1e9 / peak,
                                bytes=int(nbytes))
        del g, xt, ht, ps, counts, rowptr, col
        torch.cuda.empty_cache()
        log(f"search C2: {out['search_c2']}")
    if "dsph_4m" in which:
        # ---- plain delta-SPH at 4M
        sc = S.dam_break(dx=0.0005, block=(1.0, 1.0), scheme="dsph", seed=5)
        b = SceneBatch([sc])
        sec = _timed_steps(b, warmup, steps)
        # the step builds no fluid pair list (the cell walk); K from one
        b.search(b.x)
        nnz = int(b.ff.rowptr[b.NP].item())
        k_mean = nnz / b.NP
        cells = b.grid.ncells / b.NP
        rhs = 12 + 12 + 4 + 8 * cells + 4 + 4 * k_mean + 4 + 4 * k_mean + 24 + 12
        nbytes = b.NP * 4 * (rhs + 52)
        out["dsph_4m"] = dict(particles=b.NP, ms_per_step=sec * 1e3,
                              particle_steps_per_s=b.NP / sec, achieved_gbs=nbytes / sec / 1e9,
                              frac=nbytes / sec / 1e9 / peak, bytes_per_step=int(nbytes),
                              mean_neighbours=k_mean)
        del b
        torch.cuda.empty_cache()
        log(f"plain delta-SPH 4M: {out['dsph_4m']}")
    if "dfsph_c3" in which:
        # ---- C3: plain DFSPH at 1M
        sc = W.c3_scene()
        b = SceneBatch([sc])
        sec = _timed_steps(b, warmup, steps)
        out["dfsph_c3"] = dict(particles=b.NP, ms_per_step=sec * 1e3,
                               particle_steps_per_s=b.NP / sec,
                               iterations_last_step=int(b.dfsph_iterations()[0]))
        rl = pressure_roofline(b)
        out["dfsph_c3"]["k_pressure"] = dict(avg_launch_us=rl["seconds"] * 1e6,
                                             bytes_per_launch=rl["bytes"],
                                             achieved_gbs=rl["achieved_gbs"],
                                             frac=rl["achieved_gbs"] / peak)
        del b
        torch.cuda.empty_cache()
        log(f"plain DFSPH C3: {out['dfsph_c3']}")
    if "dsph_c1" in which:
        # ---- C1: the 5,000-particle delta-SPH dam break, one scene on the GPU
        sc = W.c1_scene()
        b = SceneBatch([sc])
        b.capture()
        for _ in range(warmup):
            b.step_graph()
        torch.cuda.synchronize()
        lat = []
        for _ in range(50):
            t0 = time.perf_counter()
            b.step_graph()
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        lat = np.array(lat)
        out["dsph_c1"] = dict(particles=b.NP, ms_per_step=float(lat.mean() * 1e3),
                              particle_steps_per_s=float(b.NP / lat.mean()),
                              p50_step_ms=float(np.percentile(lat, 50) * 1e3),
                              p99_step_ms=float(np.percentile(lat, 99) * 1e3), steps=len(lat))
        del b
        log(f"plain delta-SPH C1: {out['dsph_c1']}")
    if "dplus_4m" in which:
        out["dplus_4m"] = dplus_path()
        log(f"delta-plus 4M: {out['dplus_4m']}")
    return out


def dplus_path(reps: int = 10):
    """delta-plus surface detection and shifting on a 2000 x 2000 dam break
    (dx = 0.0005, 4,000,000 fluid particles, mDBC walls, velocities U(+-0.5)
    and densities 1 +- 0.004 so that the shift and the pressure switch act),
    through the module API: each verb timed with events over `reps` calls.
    Algorithmic bytes per fluid row (each array once; K = mean fluid
    neighbours, Kb = boundary pairs per row; columns 4 B, row pointers 4 B):
      normalization  rows/cols (4 + 4K) + x, A, rho (16) + M, L, eig (40)
      normal         rows/cols (4 + 4K) + x, V (12) + L, eig (24) + grad, n (16)
      surface        two passes: (4 + 4K) + x, n, eig (24) + mask, angles (12);
                     (4 + 4K) + mask, eig (12) + class (4)
      shift          (4 + 4K) + v (8, speed pass) + x, A, rho, n, class (28)
                     + shift (8)
      pressure       (4 + 4K) + x, rho, V, class (20) + acc read and write (16)
    plus, in every verb that reads the boundary, (4 + 4 Kb) per fluid row."""
    import torch

    from pytorchsph_amd.deltaSPH import deltaSPHSimulation

    dx = 0.0005
    cfg = {
        "particle": {"radius": dx / math.sqrt(math.pi)},
        "domain": {"min": [0.0, 0.0], "max": [2.0, 1.2]},
        "emitter": {"damBreak": {"min": [0.0, 0.0], "max": [1.0, 1.0], "velocity": [0.0, 0.0]}},
        "compute": {"device": "cuda"},
        "fluid": {"restDensity": 1000.0},
        "simulation": {"scheme": "deltaPlus", "boundaryScheme": "solid",
                       "densityScheme": "continuum"},
        "diffusion": {"densityScheme": "MOG", "kinematic": 0.001},
        "integration": {"scheme": "RK4", "dt": 1e-5},
        "shifting": {"enabled": True, "scheme": "deltaPlus"},
    }
    sim = deltaSPHSimulation(cfg)
    sim.initializeSimulation()
    s = sim.simulationState
    n = s["numParticles"]
    g = torch.Generator(device="cuda").manual_seed(7)
    s["fluidVelocity"] += (torch.rand((n, 2), device="cuda", generator=g) - 0.5)
    s["fluidDensity"] += (torch.rand(n, device="cuda", generator=g) - 0.5) * 0.008
    sim.neighborSearch.search(s, sim)
    bm = sim.boundaryModule
    bm.boundaryFilterNeighborhoods(s, sim)
    bm.boundaryNeighborhoodSearch(s, sim)
    sim.sphDensity.evaluate(s, sim)
    bm.evalBoundaryDensity(s, sim)
    sm = sim.shiftingModule
    sm.detectSurface(s, sim)
    s["fluidAcceleration"] = torch.zeros_like(s["fluidVelocity"])
    sim.pressureModule.computePressure(s, sim)
    torch.cuda.synchronize()
    K = int(sim.neighborSearch.csr.rowptr[n].item()) / n
    Kb = int(bm.fb.rowptr[n].item()) / n
    rb = 4 + 4 * Kb

    def renorm():
        s.pop("fluidL", None)
        sm.computeNormalizationmatrix(s, sim)

    verbs = {
        "normalization": (renorm, 4 + 4 * K + 16 + 40 + rb),
        "normal": (lambda: sm.computeFluidNormal(s, sim), 4 + 4 * K + 12 + 24 + 16 + rb),
        "surface": (lambda: sm.detectFluidSurface(s, sim),
                    (4 + 4 * K + 24 + 12 + rb) + (4 + 4 * K + 12 + 4)),
        "shift": (lambda: sm.computeShiftAmount(s, sim), 4 + 4 * K + 8 + 28 + 8 + rb),
        "pressure": (lambda: sim.pressureModule.computePressureAcceleration(s, sim),
                     4 + 4 * K + 20 + 16 + rb),
    }
    peak = 8000.0
    out = dict(particles=n, boundary=bm.numPtcls, mean_neighbours=K, boundary_pairs_per_row=Kb)
    total = 0.0
    for name, (fn, per_row) in verbs.items():
        sec = _events_avg(fn, torch.cuda.current_stream(), reps=reps, warm=2)
        gbs = n * per_row / sec / 1e9
        out[name] = dict(us=sec * 1e6, bytes_per_row=per_row, achieved_gbs=gbs, frac=gbs / peak)
        total += sec
    out["detect_and_shift_us"] = total * 1e6 - out["pressure"]["us"]
    surf = s["fluidSurfaceMask"]
    out["classes"] = {str(k): int((surf == k).sum().item()) for k in (0.0, 1.0, 2.0)}
    del sim, s
    torch.cuda.empty_cache()
    # the batched engine's delta-plus step (detection and the pressure switch
    # in each RK4 stage, then search, mDBC and the shift) on the 4M dam break
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd.engine import SceneBatch

    sc = S.dplus_variant(S.dam_break(dx=0.0005, block=(1.0, 1.0), scheme="dsph", seed=5))
    b = SceneBatch([sc])
    sec = _timed_steps(b, 1, 3)
    out["engine_step"] = dict(particles=b.NP, ms_per_step=sec * 1e3,
                              particle_steps_per_s=b.NP / sec)
    del b
    torch.cuda.empty_cache()
    return out


def serving_run(window_s: float):
    """Poisson arrivals (lambda 20 / s, sizes log-uniform 2k-500k, 100-2,000
    steps, schemes 50/50, default_rng(4)) over `window_s` seconds into the
    continuous-batching server; the loop stops at the end of the window and
    reports what it completed and what is still running or waiting."""
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd import workloads as W
    from pytorchsph_amd.serving import Job, Server

    t0 = time.perf_counter()
    specs = W.c5_jobs(window_s=window_s)
    jobs = [Job(id=k, scene=S.scene_for_count(j["kind"], j["scheme"], j["n"], j["seed"]),
                steps=j["steps"], t_arrive=j["t"]) for k, j in enumerate(specs)]
    log(f"serving: {len(jobs)} jobs generated in {time.perf_counter() - t0:.1f}s")
    srv = Server(log=log)
    out = srv.run(jobs, time_limit=window_s)
    out["offered_particle_steps_per_s"] = sum(j.n * j.steps for j in jobs) / window_s
    out["workload"] = (f"Poisson lambda=20/s over {window_s:g} s, 2k-500k particles, "
                       f"100-2000 steps, delta-SPH/DFSPH 50/50 (seed 4)")
    return out


def host_threads() -> int:
    """Threads for the CPU legs: the OpenMP thread count the host grants
    (OMP_NUM_THREADS), at most the cores Python sees."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def cpu_baseline(seconds: float):
    """The oracle's delta-SPH RK4 step (fp32 neighbour search, FP64 physics,
    OpenMP) on the 5,000-particle dam break, on this host's cores."""
    from oracle import oracle as O
    from pytorchsph_amd import workloads as W

    threads = host_threads()
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


def cpu_paths():
    """The oracle (fp32 search, FP64 physics, OpenMP) on the other
    single-scene workloads, on this host:
      dsph_c1_1thread   the CPU baseline's C1 step on one thread (median of
                        20 steps after 3 warm-up)
      search_c2         the 4,000,000-point search, all threads, one search
      dfsph_c3          one DFSPH step of the 1,000,000-particle dam break,
                        all threads (the search included)
      search_c2_reference   the same 4,000,000-point search by the
                        reference's own CPU code (buildNeighborList of
                        cppSrc/neighSearch.cpp, compiled from its source into
                        oracle/_ref) after its sortPositions / constructHashMap
                        steps restated in torch; all threads"""
    from oracle import oracle as O
    from oracle import ref_driver as R
    from pytorchsph_amd import workloads as W

    out = {}
    threads = host_threads()
    sc = W.c1_scene()
    so = O.SceneOracle(sc)
    O.set_threads(1)
    x, v, r = sc.x.astype(np.float64), sc.v.astype(np.float64), np.ones(sc.n)
    for _ in range(3):
        x, v, r = so.dsph_step(x, v, r)
    times = []
    for _ in range(20):
        t0 = time.perf_counter()
        x, v, r = so.dsph_step(x, v, r)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    out["dsph_c1_1thread"] = dict(particle_steps_per_s=sc.n / med, ms_per_step=med * 1e3,
                                  threads=1)
    O.set_threads(threads)
    x, h = W.c2_points()
    qx, qy, cx, _ = O.grid_params((0.0, 0.0), (1.0, 1.0), h)
    hq = np.full(len(x), h, np.float32)
    t0 = time.perf_counter()
    csr = O.search(x, hq, x, (qx, qy), h, cx, 1, self_mode=True)
    sec = time.perf_counter() - t0
    out["search_c2"] = dict(seconds=sec, searches_per_s=1.0 / sec, pairs=int(csr.nnz),
                            threads=threads)
    del csr
    if R.available():
        import torch

        torch.set_num_threads(threads)
        m = R.load()
        t0 = time.perf_counter()
        st = R.hash_structures(x, hq, (0.0, 0.0), (1.0, 1.0))
        t1 = time.perf_counter()
        rows, _ = m.buildNeighborList(st["sx"], st["sh"], st["table"], st["cell_keys"],
                                      st["counts"], st["starts"], st["sort"], st["cells_x"],
                                      st["P"], 1)
        t2 = time.perf_counter()
        out["search_c2_reference"] = dict(seconds=t2 - t0, build_list_seconds=t2 - t1,
                                          pairs=int(rows.numel()), threads=threads,
                                          kind="reference")
        del rows, st
    sc = W.c3_scene()
    so = O.SceneOracle(sc)
    t0 = time.perf_counter()
    _, _, _, res = so.dfsph_step(sc.x, sc.v)
    sec = time.perf_counter() - t0
    out["dfsph_c3"] = dict(seconds_per_step=sec, particle_steps_per_s=sc.n / sec,
                           iterations=int(res["iters"]), threads=threads)
    log(f"cpu paths: {out}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("runs on several GPUs are not implemented: scenes would shard across GPUs with "
            "no exchange; run with --gpus 1")
        sys.exit(2)
    import torch

    from pytorchsph_amd import _lib as L
    from pytorchsph_amd import workloads as W
    from pytorchsph_amd.autotune import tune
    from pytorchsph_amd.engine import MixedBatch, SceneBatch

    if args.path:
        print(json.dumps({args.path: plain_paths(args.steps, args.warmup, {args.path})[args.path]}),
              flush=True)
        return
    if args.dplus_only:
        print(json.dumps({"dplus_4m": dplus_path()}), flush=True)
        return
    t0 = time.perf_counter()
    scenes = W.c4_scenes(n_scenes=args.scenes)
    n_part = int(sum(s.n for s in scenes))
    log(f"{len(scenes)} scenes, {n_part} fluid particles, generated in "
        f"{time.perf_counter() - t0:.1f}s")
    scheme_kw = {}
    cell = None
    if not args.no_tune and not args.tuning:
        # the DFSPH part searches a list every step: try half cells
        from pytorchsph_amd.autotune import tune_cell_size
        t0 = time.perf_counter()
        cell = tune_cell_size([s for s in scenes if s.scheme == "dfsph"], log=log)
        scheme_kw["dfsph"] = {"cell_div": cell["cell_div"]}
        log(f"cell size tuned in {time.perf_counter() - t0:.1f}s")
    batch = MixedBatch(scenes, scheme_kw=scheme_kw)
    tuned = None
    if args.tuning:
        vals = [int(t) for t in args.tuning.split(",")]
        if len(vals) not in (4, 5, 6, 7):
            raise SystemExit("--tuning takes count,gather,solve[,fill[,walk[,block]]],cap")
        L.set_tuning(**dict(zip(("search", "gather", "solve", "fill", "walk", "block"),
                                vals[:-1])))
        cap = vals[-1]
        L.DFSPH_GRID_CAP = cap
        tuned = {"choice": dict(L.get_tuning(), grid_cap=cap)}
        args.no_tune = True
    elif not args.no_tune:
        t0 = time.perf_counter()
        tuned = tune(batch.parts, log=log)
        log(f"autotuned in {time.perf_counter() - t0:.1f}s: {tuned['choice']}")
    if args.roofline_only:
        dsph_part = next(b for b in batch.parts if b.scheme == "dsph")
        rl_w = cellwalk_roofline(dsph_part)
        rl_g = gather_roofline(dsph_part)
        rl = pressure_roofline(next(b for b in batch.parts if b.scheme == "dfsph"))
        copied = copy_calibration()
        rw = read_write_calibration()
        print(json.dumps({"cellwalk": rl_w, "dsph_gather": rl_g, "pressure": rl,
                          "copy_bytes": copied, **rw}), flush=True)
        return
    batch.capture()
    torch.cuda.synchronize()
    log("captured")
    # the scenes flow before the timed window (dam breaks collapse, drops
    # approach): the timed steps are not those of the fluid at rest
    for k in range(args.warmup):
        batch.step(serial=args.serial)
        if (k + 1) % 50 == 0:
            torch.cuda.synchronize()
            log(f"flow step {k + 1}")
    torch.cuda.synchronize()
    dfsph_parts = [b for b in batch.parts if b.scheme == "dfsph"]
    lat, iters = [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        batch.step(serial=args.serial)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        # the device already finished the step: reading its per-scene
        # iteration counts does not hold the GPU
        iters += [int(i) for b in dfsph_parts for i in b.dfsph_iterations()]
        log(f"step {k}: {lat[-1] * 1e3:.1f} ms")
    total = time.perf_counter() - t_start
    ms_per_step = total / args.steps * 1e3
    value = n_part * args.steps / total
    status = batch.status()

    dsph_part = next(b for b in batch.parts if b.scheme == "dsph")
    rl_w = cellwalk_roofline(dsph_part)
    log(f"delta-SPH cell walk: {rl_w}")
    rl_g = gather_roofline(dsph_part)
    log(f"delta-SPH list gather: {rl_g}")
    rl = pressure_roofline(dfsph_parts[0])
    log(f"DFSPH pressure: {rl}")
    del batch
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain:
        # the same schedule as the batch: capture (one eager step), warmup
        # steps, then the timed steps, one scene at a time
        rng = np.random.default_rng(7)
        pick = rng.choice(len(scenes), size=min(args.plain_sample, len(scenes)), replace=False)
        t_sum, n_sum = 0.0, 0
        batch_tuning = L.get_tuning()
        batch_cap = L.DFSPH_GRID_CAP
        for k in pick:
            b = SceneBatch([scenes[k]])
            if not args.no_tune:        # each scene with its own best launch shapes
                tune([b])
            b.capture()
            for _ in range(args.warmup):
                b.step_graph()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                b.step_graph()
                torch.cuda.synchronize()
            t_sum += (time.perf_counter() - t0) / args.steps
            n_sum += scenes[k].n
            del b
        plain = n_sum / t_sum
        L.set_tuning(**batch_tuning)
        L.DFSPH_GRID_CAP = batch_cap
        log(f"plain one-scene-at-a-time: {plain:.3e} particle-steps/s over {len(pick)} scenes")

    paths = None if args.no_paths else plain_paths()

    serve = None
    if args.serve_seconds > 0:
        serve = serving_run(args.serve_seconds)
        log(f"serving: {serve}")

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    cpu_more = None if (args.no_cpu or args.no_paths) else cpu_paths()
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
        "flow_steps_before_timing": args.warmup,
        "dfsph_iterations_mean": float(np.mean(iters)) if iters else None,
        # Jacobi iterations of every DFSPH scene in every timed step
        "dfsph_iterations_histogram": {str(k): int(c) for k, c in
                                       enumerate(np.bincount(iters)) if c} if iters else None,
        "status_flags": int(np.bitwise_or.reduce(status)),
        # the delta-SPH cell walk is the step's largest kernel (four launches
        # per step, each with its RK4 stage); it is bound by VALU issue, not
        # by HBM (valu_issue_frac); the list gather it replaces and the DFSPH
        # pressure update, whose count per step follows the solver's
        # iterations, are reported beside it
        "roofline": {"bound": "hbm", "achieved": rl_w["achieved_gbs"], "peak": peak,
                     "unit": "GB/s", "frac": rl_w["achieved_gbs"] / peak,
                     "traffic": traffic_from_profile(rl_w["kernel"]),
                     "traffic_source": TRAFFIC_PROFILE,
                     "kernel": rl_w["kernel"], "avg_launch_us": rl_w["seconds"] * 1e6,
                     "bytes_per_launch": rl_w["bytes"], "record_bytes": rl_w["record_bytes"],
                     "rows": rl_w["rows"], "boundary_pairs": rl_w["boundary_pairs"],
                     "valu_issue_frac": valu_issue_from_profile(rl_w["kernel"]),
                     "valu_issue_source": ISSUE_PROFILE},
        "roofline_dsph_list_gather": {"achieved": rl_g["achieved_gbs"],
                                      "frac": rl_g["achieved_gbs"] / peak,
                                      "traffic": traffic_from_profile(rl_g["kernel"]),
                                      "kernel": rl_g["kernel"],
                                      "avg_launch_us": rl_g["seconds"] * 1e6,
                                      "bytes_per_launch": rl_g["bytes"], "rows": rl_g["rows"],
                                      "pairs": rl_g["pairs"],
                                      "valu_issue_frac": valu_issue_from_profile(rl_g["kernel"])},
        "roofline_dfsph_pressure": {"achieved": rl["achieved_gbs"],
                                    "frac": rl["achieved_gbs"] / peak,
                                    "traffic": traffic_from_profile(rl["kernel"]),
                                    "kernel": rl["kernel"],
                                    "avg_launch_us": rl["seconds"] * 1e6,
                                    "bytes_per_launch": rl["bytes"], "rows": rl["rows"],
                                    "pairs": rl["pairs"],
                                    "valu_issue_frac": valu_issue_from_profile(rl["kernel"])},
        "tuning": dict(tuned["choice"], dfsph_cell_div=cell["cell_div"] if cell else 1)
                  if tuned else None,
        "cell_size_times": cell["times"] if cell else None,
        "serving": serve,
        "plain_paths": paths,
        "cpu_baseline": cpu,
        "cpu_paths": cpu_more,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
