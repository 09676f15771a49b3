"""dfsphSimulation: divergence-free SPH with the reference's module
sequence (src/dfsph.py:68-202)."""
from __future__ import annotations

import torch

from .modules.akinciBoundary import akinciBoundaryModule
from .modules.density import densityModule
from .modules.dfsph import dfsphModule
from .modules.diffusion import diffusionModule
from .modules.gravity import gravityModule
from .modules.laminar import laminarViscosityModule
from .modules.neighborSearch import neighborSearchModule
from .simulationBase import SPHSimulation


class dfsphSimulation(SPHSimulation):
    def __init__(self, config):
        super().__init__(config)
        self.neighborSearch = neighborSearchModule()
        self.sphDensity = densityModule()
        self.DFSPH = dfsphModule()
        self.gravityModule = gravityModule()
        self.modules = [self.neighborSearch, self.sphDensity, self.DFSPH, self.gravityModule]
        if self.config["diffusion"]["velocityScheme"] != "deltaSPH":
            raise NotImplementedError("only the deltaSPH velocity diffusion is on the hot path")
        self.velocityDiffusionModule = diffusionModule()
        self.laminarViscosityModule = laminarViscosityModule()
        self.modules += [self.velocityDiffusionModule, self.laminarViscosityModule]
        if self.config["simulation"]["boundaryScheme"] != "Akinci":
            raise NotImplementedError("DFSPH runs with the Akinci boundary")
        self.boundaryModule = akinciBoundaryModule()
        self.modules.append(self.boundaryModule)
        for module in self.modules:
            params = module.getParameters()
            if params is not None:
                for p in params:
                    p.parseConfig(self.config)

    def timestep(self):
        """dfsph.py:134-202."""
        s = self.simulationState
        self.neighborSearch.search(s, self)
        self.boundaryModule.boundaryFilterNeighborhoods(s, self)
        self.boundaryModule.boundaryNeighborhoodSearch(s, self)
        self.sphDensity.evaluate(s, self)
        self.boundaryModule.evalBoundaryDensity(s, self)
        s["fluidAcceleration"] = torch.zeros_like(s["fluidVelocity"])
        self.gravityModule.evaluate(s, self)
        if self.config["dfsph"]["divergenceSolver"]:
            s["divergenceIterations"] = self.DFSPH.divergenceSolver(s, self)
            s["fluidAcceleration"] += s["fluidPredAccel"]
        s["densityIterations"] = self.DFSPH.incompressibleSolver(s, self)
        s["fluidAcceleration"] += s["fluidPredAccel"]
        self.velocityDiffusionModule.evaluate(s, self)
        self.laminarViscosityModule.computeLaminarViscosity(s, self)
        continuum = self.config["simulation"]["densityScheme"] == "continuum"
        return s["fluidAcceleration"], s["fluidVelocity"], s["dpdt"] if continuum else None
