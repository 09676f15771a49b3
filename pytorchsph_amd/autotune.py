"""Launch-shape autotuner of the batched step.

The knobs are the lanes per row of the neighbour count and fill, of the
kernel-sum gathers over pair lists, of the delta-SPH cell walk and of the DFSPH
solve, the workgroup size of the list gathers and the per-query search
(sph_set_tuning), the workgroup cap of the DFSPH iteration kernels, and
(tune_cell_size, which builds the batch once per candidate) the cell size of
the grid of list batches.  Each knob is timed on the batch it will run
(same scenes, same pair lists), one knob at a time with the others held,
and the fastest value is kept.  Pair sets do not depend on any knob; sums
are re-associated, so results move within fp32 rounding.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L

LANE_VALUES = (1, 2, 4, 8, 16, 32)
BLOCK_VALUES = (64, 128, 256)
GRID_CAPS = (2048, 4096, 8192, 16384)


def _time(fn, reps: int) -> float:
    """Average seconds per call of fn() on the current stream (events)."""
    stream = torch.cuda.current_stream()
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def _search_fn(b):
    if getattr(b, "stencil_lists", False):
        return lambda: b._dfsph_search()
    return lambda: b.search(b.x)


def _gather_fn(b):
    if b.scheme == "dsph":
        return lambda: b.dsph_rhs(b.x, b.v, b.rho)
    F = b.fluid(b.x, b.v, b.rho)
    ff = b.ff.csr().c()
    fb = b.fb.csr().c() if b.NB else None
    B = b.boundary()
    sl = b.stencil_list() if b.stencil_lists else None

    def run():
        if sl is not None:
            L.call("sph_fluid_density_sl", C.byref(F), L.ptr(b.phys), C.byref(sl), C.byref(B),
                   C.byref(fb) if fb is not None else None, L.ptr(b.rho), L.ptr(b.vol),
                   L.stream_handle())
            return
        L.call("sph_fluid_density", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B),
               C.byref(fb) if fb is not None else None, b._order(), L.ptr(b.rho), L.ptr(b.vol),
               L.stream_handle())
    return run


def _solve_fn(b):
    """accel + pressure over every scene (the active list after prepare)."""
    F = b.fluid(b.x, b.v, b.rho)
    ff = b.ff.csr().c()
    fb = b.fb.csr().c() if b.NB else None
    bf = b.bf.csr().c()
    B = b.boundary()
    sh = L.stream_handle()

    sl = b.stencil_list() if b.stencil_lists else None

    def run():
        D = b._dfsph_struct()     # grid size follows the current knobs
        fbp = C.byref(fb) if fb is not None else None
        if sl is not None:
            L.call("sph_dfsph_accel_sl", C.byref(F), L.ptr(b.phys), C.byref(sl), fbp,
                   C.byref(D), sh)
            L.call("sph_dfsph_pressure_sl", C.byref(F), L.ptr(b.phys), C.byref(sl), C.byref(B),
                   fbp, C.byref(bf), C.byref(D), sh)
            return
        L.call("sph_dfsph_accel", C.byref(F), L.ptr(b.phys), C.byref(ff), fbp, C.byref(D), sh)
        L.call("sph_dfsph_pressure", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B),
               fbp, C.byref(bf), C.byref(D), sh)
    return run


def tune(parts, reps: int = 5, log=None) -> dict:
    """Tunes the knobs on the SceneBatch objects `parts` (their current
    state; DFSPH parts are left with a fresh prologue, so call this before
    stepping or capturing).  Returns {"choice": {...}, "times": {...}}."""
    dsph = [b for b in parts if b.scheme == "dsph"]
    dfsph = [b for b in parts if b.scheme == "dfsph"]
    torch.cuda.synchronize()
    for b in dsph:
        b.dsph_rhs(b.x, b.v, b.rho)
    for b in dfsph:
        b._dfsph_pre()
    torch.cuda.synchronize()
    # a knob is timed only on the parts whose kernels take it: the delta-SPH
    # right-hand side of a cell-walk batch runs on the walk knob, the list
    # gathers on the gather knob and the CSR solve on the solve knob; the
    # stencil-list DFSPH gathers run one lane per row (they keep the grid
    # cap, timed below)
    csr_dfsph = [b for b in dfsph if not getattr(b, "stencil_lists", False)]
    plans = {
        "search": [_search_fn(b) for b in parts],
        "fill": [_search_fn(b) for b in parts],
        "gather": [_gather_fn(b) for b in dsph if not b.cellwalk]
                  + [_gather_fn(b) for b in csr_dfsph],
        "walk": [_gather_fn(b) for b in dsph if b.cellwalk],
        "solve": [_solve_fn(b) for b in csr_dfsph],
    }
    # the workgroup size of the list gathers and the per-query search, timed
    # on both once their lanes are chosen
    plans["block"] = plans["search"] + plans["gather"]
    times: dict = {}
    choice = L.get_tuning()
    for knob, fns in plans.items():
        if not fns:
            continue
        res = {}
        for v in (BLOCK_VALUES if knob == "block" else LANE_VALUES):
            L.set_tuning(**{knob: v})
            res[v] = sum(_time(fn, reps) for fn in fns)
        best = min(res, key=res.get)
        L.set_tuning(**{knob: best})
        choice[knob] = best
        times[knob] = res
        if log:
            log(f"autotune {knob}: " + ", ".join(f"{k}:{t * 1e6:.0f}us" for k, t in res.items())
                + f" -> {best}")
    if dfsph:
        fns = [_solve_fn(b) for b in dfsph]
        res = {}
        for cap in GRID_CAPS:
            L.DFSPH_GRID_CAP = cap
            res[cap] = sum(_time(fn, reps) for fn in fns)
        best = min(res, key=res.get)
        L.DFSPH_GRID_CAP = best
        choice["grid_cap"] = best
        times["grid_cap"] = res
        if log:
            log("autotune grid_cap: " + ", ".join(f"{k}:{t * 1e6:.0f}us" for k, t in res.items())
                + f" -> {best}")
    # leave the DFSPH parts in a consistent prologue state
    for b in dfsph:
        b._dfsph_pre()
    torch.cuda.synchronize()
    return {"choice": choice, "times": {k: {str(a): c for a, c in v.items()}
                                        for k, v in times.items()}}


def tune_cell_size(scenes, steps: int = 3, divs=(1, 2), log=None, **kw) -> dict:
    """Cell size of a list batch of `scenes` (one scheme): each candidate
    cell (support / d, stencil radius d) gets its own batch, captured and
    timed over `steps` graph steps after one warm-up step; returns
    {"cell_div": best, "times": {d: seconds per step}}.  The pair sets are
    the same for every d; the order inside a row follows the stencil."""
    import time

    from .engine import SceneBatch

    times = {}
    for d in divs:
        b = SceneBatch(scenes, cell_div=d, **kw)
        b.capture()
        b.step_graph()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step_graph()
        torch.cuda.synchronize()
        times[d] = (time.perf_counter() - t0) / steps
        del b
        torch.cuda.empty_cache()
    best = min(times, key=times.get)
    if log:
        log("autotune cell_div: " + ", ".join(f"{d}:{s * 1e3:.2f}ms" for d, s in times.items())
            + f" -> {best}")
    return {"cell_div": best, "times": {str(d): s for d, s in times.items()}}
