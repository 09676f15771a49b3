"""momentumModule (src/modules/momentum.py:41-90): continuity equation
drho/dt_i = -sum_j ((u_j - u_i) . gradW_ij) V_j rho_j on the device, plus
the boundary module's term."""
import numpy as np

from .. import _lib as L
from ..module import Module
from ._ctx import gather


def speed_of_sound(config):
    """momentum.py:62-66: c0 = 10 sqrt(2 g 0.3) when fluid.c0 < 0."""
    c0 = 10.0 * np.sqrt(2.0 * 9.81 * 0.3)
    if config["fluid"]["c0"] < 0:
        config["fluid"]["c0"] = c0
    return config["fluid"]["c0"]


class momentumModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.restDensity = simulationConfig["fluid"]["restDensity"]
        self.c0 = speed_of_sound(simulationConfig)
        self.eps = self.support ** 2 * 0.1

    def resetState(self, simulationState):
        self.divergenceTerm = None
        self.dpdt = None
        simulationState.pop("dpdt", None)

    def computeDpDt(self, simulationState, simulation):
        self.divergenceTerm, _ = gather(simulation, simulationState, L.T_CONT)
        self.divergenceTerm += simulation.boundaryModule.computeDpDt(simulationState, simulation)
        self.dpdt = self.divergenceTerm
        simulationState["dpdt"] = self.dpdt
