"""gravityModule (src/modules/gravity.py:22-57): constant or point-source
body force, elementwise on the device."""
import torch

from ..module import Module
from ..parameter import Parameter


class gravityModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def getParameters(self):
        return [
            Parameter("gravity", "magnitude", "float", 9.81, required=False, export=True, hint=""),
            Parameter("gravity", "direction", "float array", [0, -1], required=False, export=True,
                      hint=""),
            Parameter("gravity", "pointSource", "bool", False, required=False, export=True, hint=""),
            Parameter("gravity", "potentialField", "bool", True, required=False, export=True,
                      hint=""),
            Parameter("gravity", "center", "float array", [0, 0], required=False, export=True,
                      hint=""),
        ]

    def initialize(self, simulationConfig, simulationState):
        g = simulationConfig["gravity"]
        self.magnitude = g["magnitude"]
        self.direction = g["direction"]
        self.pointSource = g["pointSource"]
        self.potentialField = g["potentialField"]
        self.center = g["center"]
        self.dtype = simulationConfig["compute"]["precision"]
        self.device = simulationConfig["compute"]["device"]

    def evaluate(self, simulationState, simulation):
        acc = simulationState["fluidAcceleration"]
        if self.pointSource:
            diff = simulationState["fluidPosition"] - torch.tensor(self.center, dtype=self.dtype,
                                                                  device=self.device)
            dist = torch.linalg.norm(diff, axis=1)
            mask = dist > 1e-7
            diff[mask] = diff[mask] / dist[mask, None]
            if self.potentialField:
                acc += -self.magnitude * diff * dist[:, None]
            else:
                acc += -self.magnitude * diff
        else:
            acc += self.magnitude * torch.tensor(self.direction, device=self.device,
                                                 dtype=self.dtype)
