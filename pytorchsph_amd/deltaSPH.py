"""deltaSPHSimulation: weakly compressible delta-SPH with the reference's
module sequence (src/deltaSPH.py:74-200)."""
from __future__ import annotations

import torch

from .modules.density import densityModule
from .modules.deltaPlus import deltaPlusModule
from .modules.densityDiffusion import densityDiffusionModule
from .modules.diffusion import diffusionModule
from .modules.gravity import gravityModule
from .modules.laminar import laminarViscosityModule
from .modules.momentum import momentumModule
from .modules.neighborSearch import neighborSearchModule
from .modules.pressure import pressureModule
from .modules.solidBoundary import solidBoundaryModule
from .simulationBase import SPHSimulation


class deltaSPHSimulation(SPHSimulation):
    def __init__(self, config):
        super().__init__(config)
        self.modules = []
        self.neighborSearch = neighborSearchModule()
        self.sphDensity = densityModule()
        self.gravityModule = gravityModule()
        self.momentumModule = momentumModule()
        self.pressureModule = pressureModule()
        self.modules += [self.neighborSearch, self.sphDensity, self.momentumModule,
                         self.pressureModule, self.gravityModule]
        if self.config["diffusion"]["velocityScheme"] != "deltaSPH":
            raise NotImplementedError("only the deltaSPH velocity diffusion is on the hot path")
        self.velocityDiffusionModule = diffusionModule()
        self.modules.append(self.velocityDiffusionModule)
        self.densityDiffusionModule = densityDiffusionModule()
        self.modules.append(self.densityDiffusionModule)
        self.laminarViscosityModule = laminarViscosityModule()
        self.modules.append(self.laminarViscosityModule)
        # deltaSPH.py:119-122
        c = self.config
        if c["simulation"]["scheme"] == "deltaPlus" or (
                c["shifting"]["scheme"] == "deltaPlus" and c["shifting"]["enabled"]):
            self.shiftingModule = deltaPlusModule()
            self.modules.append(self.shiftingModule)
        if self.config["simulation"]["boundaryScheme"] != "solid":
            raise NotImplementedError("delta-SPH runs with the solid (mDBC) boundary")
        self.boundaryModule = solidBoundaryModule()
        self.modules.append(self.boundaryModule)
        for module in self.modules:
            params = module.getParameters()
            if params is not None:
                for p in params:
                    p.parseConfig(self.config)

    def timestep(self):
        """deltaSPH.py:146-200."""
        s = self.simulationState
        self.neighborSearch.search(s, self)
        self.boundaryModule.boundaryFilterNeighborhoods(s, self)
        self.boundaryModule.boundaryNeighborhoodSearch(s, self)
        self.sphDensity.evaluate(s, self)
        self.boundaryModule.evalBoundaryDensity(s, self)
        if hasattr(self, "shiftingModule"):
            self.shiftingModule.detectSurface(s, self)
        s["fluidAcceleration"] = torch.zeros_like(s["fluidVelocity"])
        self.gravityModule.evaluate(s, self)
        self.momentumModule.computeDpDt(s, self)
        self.densityDiffusionModule.computeDensityDiffusion(s, self)
        self.pressureModule.computePressure(s, self)
        self.pressureModule.computePressureAcceleration(s, self)
        self.velocityDiffusionModule.evaluate(s, self)
        self.laminarViscosityModule.computeLaminarViscosity(s, self)
        continuum = self.config["simulation"]["densityScheme"] == "continuum"
        return s["fluidAcceleration"], s["fluidVelocity"], s["dpdt"] if continuum else None
