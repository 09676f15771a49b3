"""densityDiffusionModule (src/modules/densityDiffusion.py:125-179).

MOG: D_i = h delta c0 sum_j 2 (rho_j - rho_i)(x_j - x_i).gradW_ij /
(|x_j - x_i|^2 + eps) V_j.  deltaSPH: renormalisation matrices (fluid and
boundary neighbours) and L = pinv2x2(M), the renormalised density gradient,
then the renormalised diffusion.  Both add D to simulationState['dpdt'];
the boundary's own diffusion is not added (densityDiffusion.py:179).
"""
import ctypes as C

import torch

from .. import _lib as L
from ..module import Module
from ..parameter import Parameter
from ._ctx import byref, csr_c, fluid, gather
from .momentum import speed_of_sound


class densityDiffusionModule(Module):
    def getParameters(self):
        return [Parameter("deltaSPH", "HughesGrahamCorrection", "bool", True, required=False,
                          export=True, hint="")]

    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.restDensity = simulationConfig["fluid"]["restDensity"]
        self.delta = simulationConfig["diffusion"]["delta"]
        self.c0 = speed_of_sound(simulationConfig)
        self.eps = self.support ** 2 * 0.1
        self.scheme = simulationConfig["diffusion"]["densityScheme"]

    def resetState(self, simulationState):
        self.normalizationMatrix = None
        self.fluidL = None
        self.eigVals = None
        self.densityDiffusion = None

    def _lists(self, simulation, n):
        ns = simulation.neighborSearch
        bm = simulation.boundaryModule
        ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
        B, fb = (bm.boundary_struct(), bm.fb_struct()) if getattr(bm, "active", False) else (None, None)
        return ff, B, fb

    def computeNormalizationMatrices(self, simulationState, simulation):
        n = simulationState["numParticles"]
        dev = simulationState["fluidPosition"].device
        self.normalizationMatrix = torch.empty((n, 2, 2), dtype=torch.float32, device=dev)
        self.fluidL = torch.empty((n, 2, 2), dtype=torch.float32, device=dev)
        self.eigVals = torch.empty((n, 2), dtype=torch.float32, device=dev)
        F = fluid(simulation, simulationState)
        ff, B, fb = self._lists(simulation, n)
        L.call("sph_dsph_normalization", C.byref(F), L.ptr(simulation.phys), C.byref(ff), byref(B),
               byref(fb), None, L.ptr(self.normalizationMatrix), L.ptr(self.fluidL),
               L.ptr(self.eigVals), L.stream_handle())

    def computeRenormalizedDensityGradient(self, simulationState, simulation):
        n = simulationState["numParticles"]
        dev = simulationState["fluidPosition"].device
        self.renormalizedDensityGradient = torch.empty((n, 2), dtype=torch.float32, device=dev)
        F = fluid(simulation, simulationState)
        ff, B, fb = self._lists(simulation, n)
        L.call("sph_dsph_density_gradient", C.byref(F), L.ptr(simulation.phys), C.byref(ff),
               byref(B), byref(fb), L.ptr(self.fluidL), None,
               L.ptr(self.renormalizedDensityGradient), L.stream_handle())

    def computeDensityDiffusion(self, simulationState, simulation):
        if self.scheme == "deltaSPH":
            self.computeNormalizationMatrices(simulationState, simulation)
            self.computeRenormalizedDensityGradient(simulationState, simulation)
            self.densityDiffusion, _ = gather(simulation, simulationState, L.T_DDSPH,
                                              gradrho=self.renormalizedDensityGradient)
            simulationState["dpdt"] += self.densityDiffusion
        elif self.scheme == "MOG":
            self.densityDiffusion, _ = gather(simulation, simulationState, L.T_MOG)
            simulationState["dpdt"] += self.densityDiffusion
