"""pressureModule (src/modules/pressure.py:54-136): equation of state and
pressure acceleration a_i = -(1/rho_i) sum_j (p_i + p_j) V_j gradW_ij plus
the boundary module's term."""
import torch

from .. import _lib as L
from ..module import Module
from ..parameter import Parameter
from ._ctx import gather
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
            raise NotImplementedError("the deltaPlus pressure term is outside the hot path")
        _, self.pressureAccel = gather(simulation, simulationState, L.T_PRESS)
        self.pressureAccel += simulation.boundaryModule.computePressureAcceleration(
            simulationState, simulation)
        simulationState["fluidAcceleration"] += self.pressureAccel
