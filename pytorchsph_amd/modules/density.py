"""densityModule (src/modules/density.py:33-53): summation density
rho_i = sum_j A_j W_ij on the device, then V = A / rho."""
import ctypes as C

import torch

from .. import _lib as L
from ..module import Module
from ._ctx import csr_c, fluid


class densityModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.scheme = simulationConfig["simulation"]["densityScheme"]

    def resetState(self, simulationState):
        simulationState.pop("fluidVolume", None)

    def evaluate(self, simulationState, simulation):
        if self.scheme == "summation":
            n = simulationState["numParticles"]
            rho = torch.empty(n, dtype=torch.float32, device=simulationState["fluidArea"].device)
            F = fluid(simulation, simulationState)
            ns = simulation.neighborSearch
            ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
            L.call("sph_fluid_density", C.byref(F), L.ptr(simulation.phys), C.byref(ff), None,
                   None, None, L.ptr(rho), None, L.stream_handle())
            simulationState["fluidDensity"] = rho
        simulationState["fluidVolume"] = simulationState["fluidArea"] / simulationState["fluidDensity"]
