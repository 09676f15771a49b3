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
from ..scenes import _walls_solid, box_normals
from ..search import NeighborGrid, transpose
from ._ctx import csr_c, f32, fluid, gather
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
        self.simulationScheme = c["simulation"]["scheme"]
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
        # outward normals of the domain box (the polygon SDF gradient of
        # solidBoundary.py:158) and one body (the domain)
        self.boundaryNormals = torch.from_numpy(box_normals(bx + lo, lo, hi)).to(dev)
        self.bodyAssociation = torch.zeros(nb, dtype=torch.int64, device=dev)
        self.active = nb > 0
        self.bb = None
        self.eigVals = None

    def resetState(self, simulationState):
        self.boundaryPressure = None

    def saveState(self, perennialState, copy):
        # solidBoundary.py:266-278 / akinciBoundary.py:50-62 key set
        cl = (lambda t: torch.clone(t)) if copy else (lambda t: t)
        perennialState["boundaryPosition"] = cl(self.boundaryPositions)
        perennialState["boundaryVelocity"] = cl(self.boundaryVelocity)
        perennialState["boundaryDensity"] = cl(self.boundaryDensity)
        perennialState["boundarySupport"] = cl(self.boundarySupport)
        perennialState["boundaryRestDensity"] = cl(self.boundaryRestDensity)
        perennialState["boundaryArea"] = cl(self.boundaryVolume)
        perennialState["boundaryPressure"] = cl(self.pressure)
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
        if self.simulationScheme == "deltaPlus":
            # computePressureAccelDeltaPlus over the boundary pairs
            # (solidBoundary.py:882-887)
            s = simulationState
            acc = torch.zeros((s["numParticles"], 2), dtype=torch.float32, device=self.device)
            if self.active:
                F = fluid(simulation, s, s.get("fluidVolume"))
                B, fb = self.boundary_struct(), self.fb_struct()
                L.call("sph_dplus_pressure", C.byref(F), L.ptr(simulation.phys), None,
                       C.byref(B), C.byref(fb), L.ptr(s["fluidSurfaceMask"]), L.T_PRESS_B,
                       L.ptr(acc), L.stream_handle())
            return acc
        _, acc = gather(simulation, simulationState, L.T_PRESS_B, boundary=self)
        return acc

    def computeVelocityDiffusion(self, simulationState, simulation):
        _, acc = gather(simulation, simulationState, L.T_VISC_B, boundary=self)
        return acc

    def boundary_lists(self):
        """Static boundary->boundary pairs (solidBoundary.py:180-190), found
        once with the project's own search over the boundary points."""
        if self.bb is None:
            nb = self.numPtcls
            bx = self.boundaryPositions
            lo = bx.min(0).values.cpu().numpy()
            hi = bx.max(0).values.cpu().numpy()
            g = NeighborGrid([dict(vmin=lo, vmax=hi, hmax=self.support, p_begin=0, p_count=nb)],
                             nb, self.device)
            g.build(bx, self.bscene)
            self.bb = g.search(bx, self.boundarySupport, self.bscene, 1, self_sorted=True,
                               sort_rows=True)
        return self.bb

    def computeNormalizationMatrices(self, simulationState, simulation):
        """The boundary's own rows (solidBoundary.py:676-693): M_b over its
        fluid and boundary neighbours, self.fluidL = pinv2x2(M_b) and
        self.eigVals.  Returns the fluid rows' boundary part, which
        sph_dsph_normalization already adds, as zeros."""
        s = simulationState
        n, nb = s["numParticles"], self.numPtcls
        dev = self.device
        self.normalizationMatrix = torch.empty((nb, 2, 2), dtype=torch.float32, device=dev)
        self.fluidL = torch.empty((nb, 2, 2), dtype=torch.float32, device=dev)
        self.eigVals = torch.empty((nb, 2), dtype=torch.float32, device=dev)
        vf = s.get("fluidVolume")
        if vf is None:
            vf = s["fluidArea"] / s["fluidDensity"]
        vf = f32(vf)
        vb = f32(self.boundaryVolume / self.boundaryDensity)
        self._keep_norm = (vf, vb)
        bf = csr_c(self.bf.rowptr, self.bf.col, nb)
        bb = self.boundary_lists().c()
        L.call("sph_normalization_rows", L.ptr(self.boundaryPositions), L.ptr(self.bscene), nb,
               L.ptr(simulation.phys), C.byref(bf), L.ptr(f32(s["fluidPosition"])), L.ptr(vf),
               C.byref(bb), L.ptr(self.boundaryPositions), L.ptr(vb),
               L.ptr(self.normalizationMatrix), L.ptr(self.fluidL), L.ptr(self.eigVals),
               L.stream_handle())
        return torch.zeros((n, 2, 2), dtype=torch.float32, device=dev)

    def computeRenormalizedDensityGradient(self, simulationState, simulation):
        n = simulationState["numParticles"]
        return torch.zeros((n, 2), dtype=torch.float32, device=self.device)
