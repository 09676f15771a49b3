"""dfsphModule (src/modules/dfsph.py:103-317): the DFSPH density and
divergence solves on the device.

DFSPH(density=True) predicts v*, computes alpha and the source term (fluid
and Akinci PBSPH boundary), runs relaxed-Jacobi iterations until the FP64
mean clamped residual drops below densityThreshold (at least
minDensitySolverIterations, at most maxDensitySolverIterations + 1, the
reference's loop bounds at :205-207) and leaves the final pressure
acceleration in simulationState['fluidPredAccel'].  DFSPH(density=False) is
divergenceSolve (:236-260): the Akinci hooks contribute nothing, the source
is the velocity divergence alone, pressures are not clamped, and the loop
runs against divergenceThreshold within min/maxDivergenceSolverIterations.
The per-iteration .item() of the reference is replaced by one device flag
read per chunk of iterations.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib as L
from ..module import Module
from ..parameter import Parameter
from ._ctx import byref, csr_c, fluid


class dfsphModule(Module):
    def getParameters(self):
        return [
            Parameter("dfsph", "minDensitySolverIterations", "int", 2, required=False, export=True, hint=""),
            Parameter("dfsph", "minDivergenceSolverIterations", "int", 2, required=False, export=True, hint=""),
            Parameter("dfsph", "maxDensitySolverIterations", "int", 256, required=False, export=True, hint=""),
            Parameter("dfsph", "maxDivergenceSolverIterations", "int", 8, required=False, export=True, hint=""),
            Parameter("dfsph", "densityThreshold", "float", 1e-3, required=False, export=True, hint=""),
            Parameter("dfsph", "divergenceThreshold", "float", 1e-2, required=False, export=True, hint=""),
            Parameter("dfsph", "divergenceSolver", "bool", True, required=False, export=True, hint=""),
            Parameter("dfsph", "relaxedJacobiOmega", "float", 0.5, required=False, export=True, hint=""),
        ]

    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")
        self._buf = None
        self.chunk = 8

    def initialize(self, simulationConfig, simulationState):
        c = simulationConfig["dfsph"]
        self.support = simulationConfig["particle"]["support"]
        self.minDensitySolverIterations = c["minDensitySolverIterations"]
        self.maxDensitySolverIterations = c["maxDensitySolverIterations"]
        self.densityThreshold = c["densityThreshold"]
        self.minDivergenceSolverIterations = c["minDivergenceSolverIterations"]
        self.maxDivergenceSolverIterations = c["maxDivergenceSolverIterations"]
        self.divergenceThreshold = c["divergenceThreshold"]
        self._phys_div = None
        self.relaxedJacobiOmega = c["relaxedJacobiOmega"]
        self.backgroundPressure = simulationConfig["fluid"]["backgroundPressure"]
        if self.backgroundPressure:
            raise NotImplementedError("fluid.backgroundPressure is not part of the hot path")
        self.device = simulationConfig["compute"]["device"]

    def _buffers(self, n, nb):
        key = (n, nb)
        if self._buf is not None and self._buf[0] == key:
            return self._buf[1]
        dev = self.device
        z = lambda *s, dt=torch.float32: torch.zeros(*s, dtype=dt, device=dev)  # noqa: E731
        nb1 = max(nb, 1)
        ints = lambda a: torch.tensor(a, dtype=torch.int32, device=dev)  # noqa: E731
        b = dict(vpred=z(n, 2), aact=z(n), alpha=z(n), src=z(n), p=z(n), p2=z(n), apred=z(n, 2),
                 resid=z(n), xrec=z(n, 4), arec=z(n, 4), prec=z(n, 4), wvec=z(n, 2),
                 b_aact=z(nb1), b_alpha=z(nb1),
                 b_src=z(nb1), b_p=z(nb1), b_resid=z(nb1), bxrec=z(nb1, 4),
                 err_acc=z(L.DFSPH_ERR_SLOTS, L.dfsph_err_stride(1), dt=torch.float64), err_hist=z(1, 300, dt=torch.float64),
                 iters=z(1, dt=torch.int32), active=z(1, dt=torch.int32),
                 any_active=z(1, dt=torch.int32), status=z(1, dt=torch.int32),
                 nfluid=ints([n]), nbdy=ints([nb]), fbegin=ints([0]), bbegin=ints([0]),
                 act_scene=z(1, dt=torch.int32), act_pre=z(4, dt=torch.int32),
                 act_n=z(3, dt=torch.int32))
        self._buf = (key, b)
        return b

    def _divergence_phys(self, simulation) -> torch.Tensor:
        """The physics record with the divergence solve's threshold and
        iteration bounds (dfsph.py:240)."""
        if self._phys_div is None or self._phys_div[0] is not simulation.phys:
            rec = np.frombuffer(simulation.phys.cpu().numpy().tobytes(), dtype=L.PHYS_DTYPE).copy()
            rec["tau"][0] = self.divergenceThreshold
            rec["min_iters"][0] = self.minDivergenceSolverIterations
            rec["max_iters"][0] = self.maxDivergenceSolverIterations
            self._phys_div = (simulation.phys, L.records_to_tensor(rec, simulation.phys.device))
        return self._phys_div[1]

    def DFSPH(self, simulationState, simulation, density=True):
        n = simulationState["numParticles"]
        bm = simulation.boundaryModule
        nb = bm.numPtcls if getattr(bm, "active", False) else 0
        buf = self._buffers(n, nb)
        D = L.SphDfsph(*[L.ptr(buf[k]) for k in L.DFSPH_FIELDS], 1, 300,
                       L.dfsph_grid_blocks(n + nb), 0 if density else 1)
        F = fluid(simulation, simulationState)
        ns = simulation.neighborSearch
        ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
        # the Akinci hooks return zeros in the divergence solve
        # (akinciBoundary.py:167,190,255,319,356)
        B = bm.boundary_struct() if nb and density else None
        fb = bm.fb_struct() if nb and density else None
        bf = bm.bf_struct() if nb and density else None
        phys = simulation.phys if density else self._divergence_phys(simulation)
        sh = L.stream_handle()
        acc = simulationState["fluidAcceleration"].contiguous()
        L.call("sph_dfsph_prepare", C.byref(F), L.ptr(acc), L.ptr(phys), byref(B),
               C.byref(D), sh)
        L.call("sph_dfsph_alpha_source", C.byref(F), L.ptr(phys), C.byref(ff), byref(B),
               byref(fb), byref(bf), C.byref(D), sh)
        flag = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        while True:
            for _ in range(self.chunk):
                L.call("sph_dfsph_iterate", C.byref(F), L.ptr(phys), C.byref(ff),
                       byref(B), byref(fb), byref(bf), C.byref(D), sh)
            flag.copy_(buf["any_active"], non_blocking=True)
            torch.cuda.current_stream().synchronize()
            if int(flag[0]) == 0:
                break
        L.call("sph_dfsph_finish", C.byref(F), L.ptr(phys), C.byref(ff), byref(fb),
               C.byref(D), None, sh)
        iters = int(buf["iters"][0].item())
        errors = buf["err_hist"][0, :min(iters, 300)].cpu().numpy().tolist()
        if int(buf["status"][0].item()) & L.ST_NAN:
            raise Exception("Pressure solver became unstable!")
        if not density:
            # divergenceSolve's keys (dfsph.py:258-259, 304-309)
            simulationState["fluidPredAccel"] = buf["apred"].clone()
            simulationState["fluidPredictedVelocity"] = buf["vpred"].clone()
            simulationState["fluidActualArea"] = buf["aact"].clone()
            simulationState["fluidAlpha"] = buf["alpha"].clone()
            simulationState["fluidSourceTerm"] = buf["src"].clone()
            simulationState["fluidPressure"] = buf["p"].clone()
            simulationState["fluidPressure2"] = buf["p2"].clone()
            simulationState["residual"] = buf["resid"].clone()
            simulationState["divergenceErrors"] = errors
            simulationState["divergenceSolverPressure"] = simulationState["fluidPressure"]
            return errors
        simulationState["fluidPredAccel"] = buf["apred"].clone()
        simulationState["fluidPredictedVelocity"] = buf["vpred"].clone()
        simulationState["fluidActualArea"] = buf["aact"].clone()
        simulationState["fluidAlpha"] = buf["alpha"].clone()
        simulationState["fluidSourceTerm"] = buf["src"].clone()
        simulationState["fluidPressure"] = buf["p"].clone()
        simulationState["fluidPressure2"] = buf["p2"].clone()
        simulationState["residual"] = buf["resid"].clone()
        simulationState["densityErrors"] = errors
        simulationState["densitySolverPressure"] = simulationState["fluidPressure"]
        if nb:
            bm.boundaryPressure = buf["b_p"][:nb].clone()
        return errors

    def incompressibleSolver(self, simulationState, simulation):
        return self.DFSPH(simulationState, simulation, True)

    def divergenceSolver(self, simulationState, simulation):
        return self.DFSPH(simulationState, simulation, False)
