"""Shared plumbing of the module-level API: C-ABI structures built from a
simulation's `simulationState` tensors (one scene)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib as L


def f32(t: torch.Tensor) -> torch.Tensor:
    t = t.to(torch.float32)
    return t if t.is_contiguous() else t.contiguous()


def phys_tensor(config, device) -> tuple[torch.Tensor, dict]:
    """One SphPhysics record from a parsed config (simulationBase.py:22-111
    defaults plus the module parameters)."""
    c = config
    h = float(c["particle"]["support"])
    rec = np.zeros(1, dtype=L.PHYS_DTYPE)
    eos = {"TaitEOS": 0, "deltaSPH": 1, "compressible": 2}[c["pressure"]["fluidPressureTerm"]]
    g = c.get("gravity", {})
    direction = g.get("direction", [0, -1])
    mag = g.get("magnitude", 9.81)
    dfs = c.get("dfsph", {})
    d = dict(h=h, rho0=c["fluid"]["restDensity"], c0=c["fluid"]["c0"], gamma=c["pressure"]["gamma"],
             kappa=c["pressure"]["kappa"], alpha=c["diffusion"]["alpha"],
             delta=c["diffusion"]["delta"], nu=c["diffusion"]["kinematic"], eps=0.1 * h * h,
             gx=mag * direction[0], gy=mag * direction[1], dt=c["integration"]["dt"],
             beta=c.get("akinciBoundary", {}).get("beta", 0.125),
             omega=dfs.get("relaxedJacobiOmega", 0.5), tau=dfs.get("densityThreshold", 1e-3),
             fluid_omega=0.3, min_iters=dfs.get("minDensitySolverIterations", 2),
             max_iters=dfs.get("maxDensitySolverIterations", 256), eos=eos)
    for k, v in d.items():
        rec[0][k] = v
    return L.records_to_tensor(rec, device), d


def fluid(sim, state, vol=None) -> L.SphFluid:
    n = state["numParticles"]
    keep = [f32(state["fluidPosition"]), f32(state["fluidVelocity"]), f32(state["fluidDensity"]),
            f32(state["fluidArea"]), f32(state["fluidRestDensity"])]
    if vol is not None:
        vol = f32(vol)
        keep.append(vol)
    sim._keep = keep  # keeps converted tensors alive during the call
    return L.SphFluid(*[L.ptr(t) for t in keep[:5]], L.ptr(vol), L.ptr(sim.pscene(n)), n)


def csr_c(rowptr, col, nrows) -> L.SphCSR:
    return L.SphCSR(L.ptr(rowptr), L.ptr(col), int(nrows))


def byref(s):
    return None if s is None else C.byref(s)


def gather(sim, state, terms: int, boundary=None, gradrho=None):
    """dpdt [N] and acc [N, 2] of the requested delta-SPH terms.  Fluid
    volumes are simulationState['fluidVolume'] (density.py:53) when present."""
    n = state["numParticles"]
    dev = state["fluidPosition"].device
    dpdt = torch.zeros(n, dtype=torch.float32, device=dev)
    acc = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    F = fluid(sim, state, state.get("fluidVolume"))
    ns = sim.neighborSearch
    ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
    B, fb = (None, None)
    if boundary is not None and boundary.active:
        B, fb = boundary.boundary_struct(), boundary.fb_struct()
    L.call("sph_dsph_gather", C.byref(F), L.ptr(sim.phys), C.byref(ff), byref(B), byref(fb),
           L.ptr(gradrho), terms, None, L.ptr(dpdt), L.ptr(acc), L.stream_handle())
    return dpdt, acc
