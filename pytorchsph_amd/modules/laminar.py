"""laminarViscosityModule (src/modules/laminar.py:49-96), fluid terms only
like the reference (the boundary call is commented out at :94-95)."""
from .. import _lib as L
from ..module import Module
from ._ctx import gather


class laminarViscosityModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.kinematic = simulationConfig["diffusion"]["kinematic"]

    def resetState(self, simulationState):
        self.laminarViscosity = None

    def computeLaminarViscosity(self, simulationState, simulation):
        _, self.laminarViscosity = gather(simulation, simulationState, L.T_LAM)
        simulationState["fluidAcceleration"] += self.laminarViscosity
