"""neighborSearchModule with the reference's API
(src/modules/neighborSearch.py:174-252) on the gfx950 grid search.

search() builds the cell grid of the fluid positions and the pair list;
the list is kept as a CSR (self.csr, rows sorted, i.e. lexicographic
(i, j) order like the CUDA path's finalize sort at :153-157) and the
reference's COO view, unit vectors and q are materialised into
simulationState.  searchExisting() queries foreign points (boundary,
ghosts) against the grid of the last search(), like the reference's
stateful sortInfo / hashMap / cellMap.
"""
from __future__ import annotations

import torch

from ..module import Module
from ..parameter import Parameter
from ..search import NeighborGrid, pair_geometry
from ._ctx import f32


class neighborSearchModule(Module):
    def __init__(self):
        super().__init__("densityInterpolation", "Evaluates density at the current timestep")
        self.grid = None
        self.csr = None
        self._grid_key = None

    def getParameters(self):
        return [
            Parameter("neighborSearch", "gradientThreshold", "float", 1e-7, required=False,
                      export=True, hint=""),
            Parameter("neighborSearch", "supportScale", "float", 1.0, required=False, export=True,
                      hint=""),
            Parameter("neighborSearch", "sortNeighborhoods", "bool", True, required=False,
                      export=True, hint=""),
        ]

    def initialize(self, simulationConfig, simulationState):
        self.support = simulationConfig["particle"]["support"]
        self.maxNeighbors = simulationConfig["compute"]["maxNeighbors"]
        self.threshold = simulationConfig["neighborSearch"]["gradientThreshold"]
        self.supportScale = simulationConfig["neighborSearch"]["supportScale"]
        self.sortNeighborhoods = simulationConfig["neighborSearch"]["sortNeighborhoods"]
        self.dtype = simulationConfig["compute"]["precision"]
        self.device = simulationConfig["compute"]["device"]
        self.minDomain = simulationConfig["domain"]["virtualMin"]
        self.maxDomain = simulationConfig["domain"]["virtualMax"]

    def resetState(self, simulationState):
        simulationState.pop("fluidNeighbors", None)
        simulationState.pop("fluidDistances", None)
        simulationState.pop("fluidRadialDistances", None)

    def _grid_for(self, n, hmax):
        key = (n, float(hmax))
        if self._grid_key != key:
            self.grid = NeighborGrid([dict(vmin=self.minDomain, vmax=self.maxDomain, hmax=hmax,
                                           p_begin=0, p_count=n)], n, self.device)
            self._pscene = torch.zeros(n, dtype=torch.int32, device=self.device)
            self._grid_key = key
        return self.grid

    def search(self, simulationState, simulation):
        x = f32(simulationState["fluidPosition"])
        h = f32(simulationState["fluidSupport"])
        n = x.shape[0]
        hmax = float(h.max())                  # the reference reads float(hMax) too
        g = self._grid_for(n, hmax)
        # the reference's particle order: not cell order
        g.build(x, self._pscene, ordered=False)
        self.csr = g.search(x, h, self._pscene, 1, self_sorted=True, sort_rows=True)
        rows, cols, rhat, q = pair_geometry(self.csr, x, x, None, self.support, self.threshold)
        fluidNeighbors = torch.stack([rows, cols], dim=0)
        simulationState["fluidNeighbors"] = fluidNeighbors
        simulationState["fluidDistances"] = rhat
        simulationState["fluidRadialDistances"] = q
        return fluidNeighbors, rhat, q

    def searchExisting(self, queryPositions, querySupports, simulationState, simulation,
                       searchRadius: int = 1):
        xq = f32(queryPositions)
        hq = f32(querySupports)
        qs = torch.zeros(xq.shape[0], dtype=torch.int32, device=xq.device)
        csr = self.grid.search(xq, hq, qs, searchRadius, self_sorted=False, sort_rows=True)
        self.last_existing = csr
        rows, cols, rhat, q = pair_geometry(csr, xq, self.grid.x, hq, self.support, self.threshold)
        return torch.stack([rows, cols], dim=0), rhat, q
