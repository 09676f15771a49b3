"""deltaPlusModule (src/modules/deltaPlus.py:34-251): free-surface detection
and particle shifting for delta-plus SPH, each verb one or two HIP launches.

detectSurface (run in every right-hand side, deltaSPH.py:165-169):
  computeNormalizationmatrix  M_i over fluid and boundary neighbours,
                              L_i = pinv2x2(M_i), lambda_i = smaller
                              singular value (also the solid boundary's own
                              rows, solidBoundary.py:676-693)
  computeFluidNormal          n_i = -grad(lambda)_i / (|.| + eps)
  detectFluidSurface          angle test, one-ring spread, classes
                              0 / 2/3 / 1 / 2 in 'fluidSurfaceMask'
shift (after the integrator when shifting is enabled,
simulationBase.py:708-732):
  computeShiftAmount          -1.5 Ma 4h^2 sum (1 + 0.2 (W/W0)^4) m/(rho_i +
  adjustShiftingAmount        rho_j) gradW, clamped to 0.05 h and projected
                              on the surface; one fused launch, so
                              'shiftAmount' holds the adjusted shift after
                              either verb, as it does after the reference's
                              in-place adjustment.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib as L
from ..module import Module
from ._ctx import byref, csr_c, f32, fluid
from .momentum import speed_of_sound


class deltaPlusModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def exportState(self, simulationState, simulation, grp, mask):
        # deltaPlus.py:45-47
        if simulation.config["shifting"]["enabled"] and "fluidUpdate" in simulationState:
            grp.create_dataset("fluidShifting",
                               data=simulationState["fluidUpdate"].detach().cpu().numpy())

    def initialize(self, simulationConfig, simulationState):
        c = simulationConfig
        self.support = c["particle"]["support"]
        self.restDensity = c["fluid"]["restDensity"]
        self.c0 = speed_of_sound(c)
        self.eps = self.support ** 2 * 0.1
        self.packing = float(c["particle"]["packing"])
        self.boundaryScheme = c["simulation"]["boundaryScheme"]
        self._packing = None
        self._vmax = None

    # -------------------------------------------------------------- lists
    def _lists(self, simulationState, simulation):
        n = simulationState["numParticles"]
        ns = simulation.neighborSearch
        ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
        bm = simulation.boundaryModule
        if getattr(bm, "active", False):
            return ff, bm, bm.fb_struct()
        return ff, None, None

    # ------------------------------------------------------------- verbs
    def computeNormalizationmatrix(self, simulationState, simulation):
        s = simulationState
        if "fluidL" in s:
            self.normalizationMatrix = s["normalizationMatrix"]
            self.fluidL = s["fluidL"]
            self.eigVals = s["eigVals"]
        else:
            n = s["numParticles"]
            dev = s["fluidPosition"].device
            M = torch.empty((n, 2, 2), dtype=torch.float32, device=dev)
            Lm = torch.empty((n, 2, 2), dtype=torch.float32, device=dev)
            eig = torch.empty((n, 2), dtype=torch.float32, device=dev)
            F = fluid(simulation, s, s.get("fluidVolume"))
            ff, bm, fb = self._lists(s, simulation)
            B = bm.boundary_struct() if bm is not None else None
            L.call("sph_dsph_normalization", C.byref(F), L.ptr(simulation.phys), C.byref(ff),
                   byref(B), byref(fb), None, L.ptr(M), L.ptr(Lm), L.ptr(eig), L.stream_handle())
            if bm is not None:
                # the boundary's own rows; the fluid rows' boundary part is
                # already inside sph_dsph_normalization
                M += bm.computeNormalizationMatrices(s, simulation)
            s["normalizationMatrix"], s["fluidL"], s["eigVals"] = M, Lm, eig
        s["fluidLambda"] = s["eigVals"][:, 1]

    def computeFluidNormal(self, simulationState, simulation):
        s = simulationState
        n = s["numParticles"]
        dev = s["fluidPosition"].device
        grad = torch.empty((n, 2), dtype=torch.float32, device=dev)
        normal = torch.empty((n, 2), dtype=torch.float32, device=dev)
        F = fluid(simulation, s, s.get("fluidVolume"))
        ff, bm, fb = self._lists(s, simulation)
        bx = beig = bvol = None
        if bm is not None:
            bx, beig = bm.boundaryPositions, bm.eigVals
            bvol = self._boundary_volumes_quirk(s, bm)
        L.call("sph_dplus_normal", C.byref(F), L.ptr(simulation.phys), C.byref(ff),
               L.ptr(s["fluidL"]), L.ptr(s["eigVals"]), L.ptr(bx), L.ptr(beig), L.ptr(bvol),
               byref(fb), L.ptr(grad), L.ptr(normal), L.stream_handle())
        s["lambdaGrad"] = grad
        s["fluidNormal"] = normal

    @staticmethod
    def _boundary_volumes_quirk(s, bm):
        """deltaPlus.py:151 takes the volume of boundary neighbour b from the
        fluid arrays at index b.  The same values, per boundary index; a pair
        whose boundary index has no fluid particle raises as the reference's
        indexing does."""
        n, nb = s["numParticles"], bm.numPtcls
        vol = torch.zeros(nb, dtype=torch.float32, device=s["fluidPosition"].device)
        m = min(n, nb)
        vol[:m] = f32(s["fluidArea"][:m]) / f32(s["fluidDensity"][:m])
        if nb > n:
            nnz = int(bm.fb.rowptr[-1].item())
            if nnz and bool((bm.fb.col[:nnz] >= n).any()):
                raise IndexError("deltaPlus: boundary neighbour index beyond the fluid arrays "
                                 "(deltaPlus.py:151)")
        return vol

    def detectFluidSurface(self, simulationState, simulation):
        s = simulationState
        n = s["numParticles"]
        dev = s["fluidPosition"].device
        tmp = torch.empty(n, dtype=torch.float32, device=dev)
        amin = torch.empty(n, dtype=torch.float32, device=dev)
        amax = torch.empty(n, dtype=torch.float32, device=dev)
        surf = torch.empty(n, dtype=torch.float32, device=dev)
        F = fluid(simulation, s)
        ff, bm, fb = self._lists(s, simulation)
        bx = bm.boundaryPositions if bm is not None else None
        L.call("sph_dplus_surface", C.byref(F), L.ptr(simulation.phys), C.byref(ff),
               L.ptr(s["fluidNormal"]), L.ptr(s["eigVals"]), L.ptr(bx), byref(fb), L.ptr(tmp),
               L.ptr(amin), L.ptr(amax), L.ptr(surf), L.stream_handle())
        s["angleMin"], s["angleMax"] = amin, amax
        s["fluidSurfaceMask"] = surf

    def computeShiftAmount(self, simulationState, simulation):
        s = simulationState
        n = s["numParticles"]
        dev = s["fluidPosition"].device
        if self._packing is None:
            self._packing = torch.tensor([self.packing], dtype=torch.float32, device=dev)
            self._vmax = torch.empty(1, dtype=torch.int32, device=dev)
        shift = torch.empty((n, 2), dtype=torch.float32, device=dev)
        F = fluid(simulation, s)
        ff, bm, fb = self._lists(s, simulation)
        B = bm.boundary_struct() if bm is not None else None
        L.call("sph_dplus_shift", C.byref(F), L.ptr(simulation.phys), 1, L.ptr(self._packing),
               C.byref(ff), byref(B), byref(fb), L.ptr(s["fluidNormal"]),
               L.ptr(s["fluidSurfaceMask"]), L.ptr(self._vmax), L.ptr(shift), L.stream_handle())
        s["shiftAmount"] = shift

    def adjustShiftingAmount(self, simulationState, simulation):
        # the clamp and the surface projection ran inside sph_dplus_shift
        simulationState["fluidUpdate"] = simulationState["shiftAmount"]

    def detectSurface(self, simulationState, simulation):
        self.computeNormalizationmatrix(simulationState, simulation)
        self.computeFluidNormal(simulationState, simulation)
        self.detectFluidSurface(simulationState, simulation)

    def shift(self, simulationState, simulation):
        if "fluidSurfaceMask" not in simulationState:
            self.detectSurface(simulationState, simulation)
        self.computeShiftAmount(simulationState, simulation)
        self.adjustShiftingAmount(simulationState, simulation)
