"""diffusionModule (src/modules/diffusion.py:47-93): delta-SPH artificial
viscosity h alpha c0 sum_j pi_ij V_j rho_j / (rho_i + rho_j) gradW_ij with
pi_ij = min(0, u_ji . x_ji) / (|x_ji|^2 + eps), plus the boundary term."""
from .. import _lib as L
from ..module import Module
from ._ctx import gather
from .momentum import speed_of_sound


class diffusionModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.alpha = simulationConfig["diffusion"]["alpha"]
        self.boundaryDiffusion = simulationConfig["diffusion"]["boundaryDiffusion"]
        self.c0 = speed_of_sound(simulationConfig)

    def resetState(self, simulationState):
        self.velocityDiffusion = None

    def evaluate(self, simulationState, simulation):
        _, self.velocityDiffusion = gather(simulation, simulationState, L.T_VISC)
        if self.boundaryDiffusion:
            self.velocityDiffusion += simulation.boundaryModule.computeVelocityDiffusion(
                simulationState, simulation)
        simulationState["fluidAcceleration"] += self.velocityDiffusion
