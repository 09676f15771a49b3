"""akinciBoundaryModule: one row of Akinci boundary particles for DFSPH
(src/modules/akinciBoundary.py).

Volumes V_b = gamma / sum_b' W_bb' and the density term sum_b' W_bb' V_b'
are set once (:137-157).  Per step: boundary->fluid pairs against the fluid
grid (:426-433), the fluid's boundary density contribution and the boundary
density rho_b = term + sum_f A_f W + beta (:395-420).  The PBSPH pressure
hooks of the DFSPH solve (:166-389) run fused inside the sph_dfsph_* kernels,
which read this module's arrays through boundary_struct() / bf / fb.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib as L
from ..module import BoundaryModule
from ..parameter import Parameter
from ..scenes import _walls_akinci, akinci_volumes, box_normals
from ..search import transpose
from ._ctx import csr_c, fluid


class akinciBoundaryModule(BoundaryModule):
    def getParameters(self):
        return [
            Parameter("akinciBoundary", "recomputeBoundary", "bool", False, required=False,
                      export=True, hint=""),
            Parameter("akinciBoundary", "beta", "float", 0.125, required=False, export=True,
                      hint=""),
        ]

    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")
        self.active = False
        self.boundaryToFluidNeighbors = None

    def initialize(self, simulationConfig, simulationState):
        c = simulationConfig
        self.support = c["particle"]["support"]
        self.beta = c["akinciBoundary"]["beta"]
        self.gamma = c["akinciBoundary"]["gamma"]
        self.device = c["compute"]["device"]
        self.pressureScheme = c["pressure"]["boundaryPressureTerm"]
        if self.pressureScheme != "PBSPH":
            raise NotImplementedError(f"Akinci boundary pressure {self.pressureScheme}: only PBSPH")
        dx = float(np.sqrt(c["particle"]["area"]))
        lo = np.array(c["domain"]["min"], dtype=np.float64)
        hi = np.array(c["domain"]["max"], dtype=np.float64)
        bx = _walls_akinci(hi[0] - lo[0], hi[1] - lo[1], dx) + lo
        vol, term = akinci_volumes(bx.astype(np.float32), self.support, self.gamma)
        dev = self.device
        self.boundaryPositions = torch.tensor(bx, dtype=torch.float32, device=dev)
        self.numPtcls = nb = self.boundaryPositions.shape[0]
        self.boundaryVolume = torch.from_numpy(vol).to(dev)
        self.boundaryDensityTerm = torch.from_numpy(term).to(dev)
        self.boundaryDensity = self.boundaryDensityTerm.clone()
        self.boundarySupport = torch.full((nb,), self.support, dtype=torch.float32, device=dev)
        self.boundaryRestDensity = torch.full((nb,), c["fluid"]["restDensity"], dtype=torch.float32,
                                              device=dev)
        self.boundaryVelocity = torch.zeros((nb, 2), dtype=torch.float32, device=dev)
        self.boundaryPressure = torch.zeros(nb, dtype=torch.float32, device=dev)
        self.bscene = torch.zeros(nb, dtype=torch.int32, device=dev)
        self.boundaryNormals = torch.from_numpy(box_normals(bx, lo, hi)).to(dev)
        self.bodyAssociation = torch.zeros(nb, dtype=torch.int64, device=dev)
        self.active = nb > 0

    def saveState(self, perennialState, copy):
        # solidBoundary.py:266-278 / akinciBoundary.py:50-62 key set
        cl = (lambda t: torch.clone(t)) if copy else (lambda t: t)
        perennialState["boundaryPosition"] = cl(self.boundaryPositions)
        perennialState["boundaryVelocity"] = cl(self.boundaryVelocity)
        perennialState["boundaryDensity"] = cl(self.boundaryDensity)
        perennialState["boundarySupport"] = cl(self.boundarySupport)
        perennialState["boundaryRestDensity"] = cl(self.boundaryRestDensity)
        perennialState["boundaryArea"] = cl(self.boundaryVolume)
        perennialState["boundaryPressure"] = cl(self.boundaryPressure)
        perennialState["boundaryNormals"] = cl(self.boundaryNormals)
        perennialState["boundaryBodyAssociation"] = cl(self.bodyAssociation)
        perennialState["boundaryParticles"] = self.numPtcls

    def exportState(self, simulationState, simulation, grp, mask):
        # solidBoundary.py:253-261: per-step boundary only when it is not static
        if not simulation.config["export"]["staticBoundary"]:
            for key, t in (("boundaryPosition", self.boundaryPositions),
                           ("boundaryVelocity", self.boundaryVelocity),
                           ("boundarySupport", self.boundarySupport),
                           ("boundaryRestDensity", self.boundaryRestDensity),
                           ("boundaryArea", self.boundaryVolume),
                           ("boundaryNormals", self.boundaryNormals),
                           ("boundaryBodyAssociation", self.bodyAssociation)):
                grp.create_dataset(key, data=t.detach().cpu().numpy())

    def boundaryNeighborhoodSearch(self, simulationState, simulation):
        if not self.active:
            return
        ns = simulation.neighborSearch
        (self.boundaryToFluidNeighbors, self.boundaryToFluidNeighborDistances,
         self.boundaryToFluidNeighborRadialDistances) = ns.searchExisting(
            self.boundaryPositions, self.boundarySupport, simulationState, simulation)
        self.bf = ns.last_existing
        self.fb = transpose(self.bf, simulationState["numParticles"])

    def boundary_struct(self) -> L.SphBoundary:
        return L.SphBoundary(L.ptr(self.boundaryPositions), L.ptr(self.boundaryVelocity),
                             L.ptr(self.boundaryDensity), L.ptr(self.boundaryVolume),
                             L.ptr(self.boundaryPressure), L.ptr(self.boundaryRestDensity),
                             L.ptr(self.bscene), self.numPtcls)

    def fb_struct(self) -> L.SphCSR:
        return csr_c(self.fb.rowptr, self.fb.col, self.fb.nrows)

    def bf_struct(self) -> L.SphCSR:
        return csr_c(self.bf.rowptr, self.bf.col, self.numPtcls)

    def evalBoundaryDensity(self, simulationState, simulation):
        n = simulationState["numParticles"]
        dev = simulationState["fluidPosition"].device
        if self.boundaryToFluidNeighbors is None:
            return torch.zeros(n, dtype=torch.float32, device=dev)
        # fluid contribution sum_b V_b W_ib: the density gather over an empty
        # fluid list and the boundary list
        empty = torch.zeros(n + 1, dtype=torch.int32, device=dev)
        F = fluid(simulation, simulationState)
        ff0 = csr_c(empty, empty, n)
        contrib = torch.empty(n, dtype=torch.float32, device=dev)
        B = self.boundary_struct()
        fb = self.fb_struct()
        L.call("sph_fluid_density", C.byref(F), L.ptr(simulation.phys), C.byref(ff0), C.byref(B),
               C.byref(fb), None, L.ptr(contrib), None, L.stream_handle())
        bf = self.bf_struct()
        rho_b = torch.empty(self.numPtcls, dtype=torch.float32, device=dev)
        L.call("sph_akinci_boundary_density", C.byref(F), L.ptr(simulation.phys),
               L.ptr(self.boundaryPositions), L.ptr(self.bscene), self.numPtcls,
               L.ptr(self.boundaryDensityTerm), C.byref(bf), L.ptr(rho_b), L.stream_handle())
        self.boundaryDensityContribution = contrib
        self.boundaryDensity = rho_b
        simulationState["fluidDensity"] += contrib

    def computeVelocityDiffusion(self, simulationState, simulation):
        return torch.zeros_like(simulationState["fluidAcceleration"])

    def dfsphPrepareSolver(self, simulationState, simulation, density=True):
        return   # folded into sph_dfsph_prepare
