"""pressureModule (src/modules/pressure.py:54-136): equation of state and
pressure acceleration a_i = -(1/rho_i) sum_j (p_i + p_j) V_j gradW_ij plus
the boundary module's term; the delta-plus scheme switches to p_j - p_i for
rows with negative pressure in the bulk class (pressure.py:40-52)."""
import ctypes as C

import torch

from .. import _lib as L
from ..module import Module
from ..parameter import Parameter
from ._ctx import csr_c, fluid, gather
from .momentum import speed_of_sound


def equation_of_state(scheme, rho_hat, rho0, c0, gamma, kappa):
    """pressure.py:101-106 (Tait, linear delta-SPH, compressible)."""
    if scheme == "TaitEOS":
        return rho0 * c0 ** 2 / gamma * (torch.pow(rho_hat, gamma) - 1)
    if scheme == "deltaSPH":
        return c0 ** 2 * (rho_hat * rho0 - rho0)
    return kappa * (rho_hat * rho0 - rho0)


class pressureModule(Module):
    def getParameters(self):
        return [Parameter("pressure", "HughesGrahamCorrection", "bool", True, required=False,
                          export=True, hint="")]

    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.restDensity = simulationConfig["fluid"]["restDensity"]
        self.gamma = simulationConfig["pressure"]["gamma"]
        self.kappa = simulationConfig["pressure"]["kappa"]
        self.fluidPressureScheme = simulationConfig["pressure"]["fluidPressureTerm"]
        self.simulationScheme = simulationConfig["simulation"]["scheme"]
        self.c0 = speed_of_sound(simulationConfig)

    def resetState(self, simulationState):
        self.pressure = None
        self.pressureAccel = None

    def computePressure(self, simulationState, simulation):
        self.pressure = equation_of_state(self.fluidPressureScheme, simulationState["fluidDensity"],
                                          self.restDensity, self.c0, self.gamma, self.kappa)
        simulation.boundaryModule.computePressure(simulationState, simulation)
        simulationState["fluidPressure"] = self.pressure
        return self.pressure

    def computePressureAcceleration(self, simulationState, simulation):
        if self.simulationScheme == "deltaPlus":
            # computePressureAccelDeltaPlus (pressure.py:40-52): p_j - p_i
            # for a row with negative pressure in the bulk class
            s = simulationState
            n = s["numParticles"]
            self.pressureAccel = torch.zeros((n, 2), dtype=torch.float32,
                                             device=s["fluidPosition"].device)
            F = fluid(simulation, s, s.get("fluidVolume"))
            ns = simulation.neighborSearch
            ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
            L.call("sph_dplus_pressure", C.byref(F), L.ptr(simulation.phys), C.byref(ff), None,
                   None, L.ptr(s["fluidSurfaceMask"]), L.T_PRESS, L.ptr(self.pressureAccel),
                   L.stream_handle())
        else:
            _, self.pressureAccel = gather(simulation, simulationState, L.T_PRESS)
        self.pressureAccel += simulation.boundaryModule.computePressureAcceleration(
            simulationState, simulation)
        simulationState["fluidAcceleration"] += self.pressureAccel
