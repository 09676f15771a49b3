"""solidBoundaryModule: ghost-particle (mDBC) walls for delta-SPH
(src/modules/solidBoundary.py).

Boundary particles sit in `layers` rows outside the domain walls with ghost
particles mirrored into the fluid (solidBoundary.py:139-161).  Per right-hand
side: boundary->fluid and ghost->fluid pairs against the fluid grid
(:618-642), the mDBC density extrapolation (:561-616) and the fluid terms of
the boundary (continuity :746-769, EOS :772-816, pressure :846-895,
velocity diffusion :820-844), each a HIP gather over the fluid-row
transpose of the boundary pairs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib as L
from ..module import BoundaryModule
from ..parameter import Parameter
from ..scenes import _walls_solid
from ..search import transpose
from ._ctx import csr_c, fluid, gather
from .pressure import equation_of_state

MLSscale = 1


class solidBoundaryModule(BoundaryModule):
    def getParameters(self):
        return [
            Parameter("solidBoundary", "recomputeBoundary", "bool", False, required=False,
                      export=True, hint=""),
            Parameter("solidBoundary", "layers", "int", 3, required=False, export=True, hint=""),
        ]

    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")
        self.active = False

    def initialize(self, simulationConfig, simulationState):
        c = simulationConfig
        self.support = c["particle"]["support"]
        self.layers = c["solidBoundary"]["layers"]
        self.device = c["compute"]["device"]
        self.restDensity = c["fluid"]["restDensity"]
        self.boundaryPressureScheme = c["pressure"]["boundaryPressureTerm"]
        self.fluidPressureScheme = c["pressure"]["fluidPressureTerm"]
        self.gamma = c["pressure"]["gamma"]
        self.kappa = c["pressure"]["kappa"]
        self.c0 = c["fluid"]["c0"]
        self.simulationDensityScheme = c["simulation"]["densityScheme"]
        dx = float(np.sqrt(c["particle"]["area"]))
        lo = np.array(c["domain"]["min"], dtype=np.float64)
        hi = np.array(c["domain"]["max"], dtype=np.float64)
        bx, gx = _walls_solid(hi[0] - lo[0], hi[1] - lo[1], dx, self.layers)
        dev = self.device
        self.boundaryPositions = torch.tensor(bx + lo, dtype=torch.float32, device=dev)
        self.boundaryGhostPositions = torch.tensor(gx + lo, dtype=torch.float32, device=dev)
        self.numPtcls = nb = self.boundaryPositions.shape[0]
        self.boundaryVolume = torch.full((nb,), np.pi * c["particle"]["radius"] ** 2,
                                         dtype=torch.float32, device=dev)
        self.boundarySupport = torch.full((nb,), self.support, dtype=torch.float32, device=dev)
        self.boundaryRestDensity = torch.full((nb,), self.restDensity, dtype=torch.float32,
                                              device=dev)
        self.boundaryVelocity = torch.zeros((nb, 2), dtype=torch.float32, device=dev)
        self.boundaryDensity = torch.ones(nb, dtype=torch.float32, device=dev)
        self.pressure = torch.zeros(nb, dtype=torch.float32, device=dev)
        self.bscene = torch.zeros(nb, dtype=torch.int32, device=dev)
        self.active = nb > 0

    def resetState(self, simulationState):
        self.boundaryPressure = None

    def saveState(self, perennialState, copy):
        perennialState["boundaryPosition"] = self.boundaryPositions
        perennialState["boundaryDensity"] = self.boundaryDensity
        perennialState["boundaryParticles"] = self.numPtcls

    # ------------------------------------------------------------ search
    def boundaryNeighborhoodSearch(self, simulationState, simulation):
        if not self.active:
            return
        ns = simulation.neighborSearch
        (self.boundaryToFluidNeighbors, self.boundaryToFluidNeighborDistances,
         self.boundaryToFluidNeighborRadialDistances) = ns.searchExisting(
            self.boundaryPositions, self.boundarySupport, simulationState, simulation)
        self.bf = ns.last_existing
        (self.ghostToFluidNeighbors, self.ghostToFluidNeighborDistances,
         self.ghostToFluidNeighborRadialDistances) = ns.searchExisting(
            self.boundaryGhostPositions, self.boundarySupport * MLSscale, simulationState,
            simulation, searchRadius=MLSscale)
        self.gf = ns.last_existing
        self.fb = transpose(self.bf, simulationState["numParticles"])

    def boundary_struct(self) -> L.SphBoundary:
        return L.SphBoundary(L.ptr(self.boundaryPositions), L.ptr(self.boundaryVelocity),
                             L.ptr(self.boundaryDensity), L.ptr(self.boundaryVolume),
                             L.ptr(self.pressure), L.ptr(self.boundaryRestDensity),
                             L.ptr(self.bscene), self.numPtcls)

    def fb_struct(self) -> L.SphCSR:
        return csr_c(self.fb.rowptr, self.fb.col, self.fb.nrows)

    # ----------------------------------------------------------- density
    def evalBoundaryDensity(self, simulationState, simulation):
        if self.simulationDensityScheme == "continuum":
            self.computeBoundaryDensitymDBC(simulationState, simulation)
        else:
            raise NotImplementedError("summation density with solid walls is outside the hot path")

    def computeBoundaryDensitymDBC(self, simulationState, simulation):
        F = fluid(simulation, simulationState)
        gf = csr_c(self.gf.rowptr, self.gf.col, self.numPtcls)
        self.neighCounts = torch.empty(self.numPtcls, dtype=torch.int32, device=self.device)
        L.call("sph_mdbc_density", C.byref(F), L.ptr(simulation.phys),
               L.ptr(self.boundaryPositions), L.ptr(self.boundaryGhostPositions),
               L.ptr(self.bscene), self.numPtcls, C.byref(gf), L.ptr(self.boundaryDensity), None,
               L.ptr(self.neighCounts), L.stream_handle())
        self.fluidVolume = self.boundaryVolume / self.boundaryDensity

    # ------------------------------------------------------- fluid terms
    def computeDpDt(self, simulationState, simulation):
        dpdt, _ = gather(simulation, simulationState, L.T_CONT_B, boundary=self)
        return dpdt

    def computePressure(self, simulationState, simulation):
        if self.boundaryPressureScheme not in ("PBSPH", "EOS"):
            raise NotImplementedError(f"boundary pressure {self.boundaryPressureScheme}")
        self.pressure = equation_of_state(self.fluidPressureScheme, self.boundaryDensity,
                                          self.restDensity, self.c0, self.gamma, self.kappa)

    def computePressureAcceleration(self, simulationState, simulation):
        _, acc = gather(simulation, simulationState, L.T_PRESS_B, boundary=self)
        return acc

    def computeVelocityDiffusion(self, simulationState, simulation):
        _, acc = gather(simulation, simulationState, L.T_VISC_B, boundary=self)
        return acc

    def computeNormalizationMatrices(self, simulationState, simulation):
        # the fluid rows' boundary part is added inside sph_dsph_normalization
        n = simulationState["numParticles"]
        return torch.zeros((n, 2, 2), dtype=torch.float32, device=self.device)

    def computeRenormalizedDensityGradient(self, simulationState, simulation):
        n = simulationState["numParticles"]
        return torch.zeros((n, 2), dtype=torch.float32, device=self.device)
