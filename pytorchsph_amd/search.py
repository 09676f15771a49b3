"""Device neighbour search: grid construction and CSR pair lists.

`grid_params` restates the reference's grid set-up arithmetic in fp32
(src/modules/neighborSearch.py:34-47): qMin = minDomain - hMax,
qMax = maxDomain + 2 hMax, cellsX = ceil((qMax - qMin) / hMax).
`NeighborGrid` owns the device buffers of one grid over many scenes and runs
sph_grid_build / sph_neighbor_count / sph_scan_exclusive / sph_neighbor_fill.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L


def grid_params(vmin, vmax, hmax):
    """Returns (qmin_x, qmin_y, cells_x, cells_y, hmax) as fp32 / int32 values,
    computed exactly as torch does it in fp32 (neighborSearch.py:34-47)."""
    f32 = np.float32
    h = f32(hmax)
    vmin = np.asarray(vmin, dtype=np.float64).astype(f32)
    vmax = np.asarray(vmax, dtype=np.float64).astype(f32)
    qmin = (vmin - h * f32(1.0)).astype(f32)
    qmax = (vmax + f32(2) * h * f32(1.0)).astype(f32)
    ext = (qmax - qmin).astype(f32)
    cells = np.ceil((ext / (h * f32(1.0))).astype(f32)).astype(np.int32)
    return f32(qmin[0]), f32(qmin[1]), int(cells[0]), int(cells[1]), h


@dataclass
class CSR:
    """Pair list in CSR form: row r holds col[rowptr[r]:rowptr[r+1]]."""
    rowptr: torch.Tensor   # int32 [nrows + 1]
    col: torch.Tensor      # int32 [capacity]
    nrows: int
    nnz: int | None = None  # known after a host read, else None

    def c(self) -> L.SphCSR:
        return L.csr(self.rowptr, self.col, self.nrows)


class NeighborGrid:
    """Cell grid over a point set that may span many scenes.

    scenes: list of dicts with keys vmin, vmax (2-sequences), hmax (float),
    p_begin, p_count.  Points of scene s are x[p_begin:p_begin+p_count].
    """

    def __init__(self, scenes, np_total: int, device="cuda"):
        self.device = torch.device(device)
        desc = np.zeros(len(scenes), dtype=L.GRID_DESC_DTYPE)
        cell_begin = 0
        self.scene_params = []
        for s, sc in enumerate(scenes):
            qx, qy, cx, cy, h = grid_params(sc["vmin"], sc["vmax"], sc["hmax"])
            desc[s] = (qx, qy, h, cx, cy, cell_begin, sc["p_begin"], sc["p_count"])
            self.scene_params.append((qx, qy, cx, cy, h))
            cell_begin += cx * cy + 1     # + the scene's bucket of outside points
        if cell_begin >= 2**31 - 1:
            raise ValueError("cell table exceeds int32 range")
        self.ncells = cell_begin
        self.nscenes = len(scenes)
        self.np = int(np_total)
        dev = self.device
        self.desc_host = desc
        self.desc = L.records_to_tensor(desc, dev)
        n1 = max(self.np, 1)
        self.cell_count = torch.empty(self.ncells, dtype=torch.int32, device=dev)
        self.cell_start = torch.empty(self.ncells + 1, dtype=torch.int32, device=dev)
        self.key = torch.empty(n1, dtype=torch.int32, device=dev)
        self.slot = torch.empty(n1, dtype=torch.int32, device=dev)
        self.sorted_idx = torch.empty(n1, dtype=torch.int32, device=dev)
        self.sorted_x = torch.empty((n1, 2), dtype=torch.float32, device=dev)
        self.status = torch.zeros(self.nscenes, dtype=torch.int32, device=dev)
        lib = L.load()
        self.ws_bytes = int(lib.sph_grid_workspace_bytes(self.ncells))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.pscene = None
        self.x = None

    def _grid(self) -> L.SphGrid:
        return L.SphGrid(
            L.ptr(self.desc), L.ptr(self.pscene), L.ptr(self.x), self.np, self.nscenes,
            self.ncells, L.ptr(self.cell_count), L.ptr(self.cell_start), L.ptr(self.key),
            L.ptr(self.slot), L.ptr(self.sorted_idx), L.ptr(self.sorted_x), L.ptr(self.status))

    def build(self, x: torch.Tensor, pscene: torch.Tensor, stream=None,
              ordered: bool = True) -> None:
        """Sorts the points into the cell table (positions x [np, 2] fp32).
        ordered: the points are (nearly) in cell order, as the engine keeps
        its state, and the per-cell atomics are cheap (sph_grid_build);
        False: any order, by the radix sort (sph_grid_build_sorted).  The
        result is the same."""
        assert x.dtype == torch.float32 and x.is_contiguous() and x.shape == (self.np, 2)
        assert pscene.dtype == torch.int32 and pscene.numel() == self.np
        self.x = x
        self.pscene = pscene
        g = self._grid()
        if ordered:
            L.call("sph_grid_build", C.byref(g), L.ptr(self.ws), self.ws_bytes,
                   L.stream_handle(stream))
            return
        ws = self._sort_ws()
        L.call("sph_grid_build_sorted", C.byref(g), None, L.ptr(ws), int(ws.numel()),
               L.stream_handle(stream))

    def _sort_ws(self) -> torch.Tensor:
        """Workspace of the radix-sort build (allocated on first use)."""
        if getattr(self, "_sws", None) is None:
            n = int(L.load().sph_grid_sort_workspace_bytes(self.np, self.ncells))
            self._sws = torch.empty(n, dtype=torch.uint8, device=self.device)
        return self._sws

    def query_struct(self, xq, hq, qscene, radius=1, self_sorted=False, qcell=None) -> L.SphQuery:
        return L.SphQuery(L.ptr(xq), L.ptr(hq), L.ptr(qscene), int(xq.shape[0]), int(radius),
                          1 if self_sorted else 0, L.ptr(qcell))

    def count(self, q: L.SphQuery, counts: torch.Tensor, stream=None) -> None:
        g = self._grid()
        L.call("sph_neighbor_count", C.byref(g), C.byref(q), L.ptr(counts), L.stream_handle(stream))

    def fill(self, q: L.SphQuery, rowptr: torch.Tensor, col: torch.Tensor, stream=None) -> None:
        g = self._grid()
        L.call("sph_neighbor_fill", C.byref(g), C.byref(q), L.ptr(rowptr), L.ptr(col),
               int(col.numel()), L.stream_handle(stream))

    def search(self, xq, hq, qscene, radius=1, self_sorted=False, sort_rows=False,
               stream=None) -> CSR:
        """Eager search with exact allocation (one host read of nnz, like the
        reference's .item() at neighSearch_cuda.cu:267)."""
        nq = int(xq.shape[0])
        q = self.query_struct(xq, hq, qscene, radius, self_sorted)
        dev = self.device
        counts = torch.empty(max(nq, 1), dtype=torch.int32, device=dev)
        rowptr = torch.empty(nq + 1, dtype=torch.int32, device=dev)
        self.count(q, counts, stream)
        scan(counts, nq, rowptr, stream)
        nnz = int(rowptr[nq].item())
        col = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        self.fill(q, rowptr, col, stream)
        out = CSR(rowptr, col, nq, nnz)
        if sort_rows:
            sort_rows_(out, stream)
        return out


_scan_ws: dict = {}


def scan(counts: torch.Tensor, n: int, out: torch.Tensor, stream=None, ws: torch.Tensor | None = None):
    """out[0] = 0, out[k+1] = sum(counts[:k+1])."""
    lib = L.load()
    need = int(lib.sph_scan_workspace_bytes(n))
    if ws is None:
        key = (counts.device, need)
        ws = _scan_ws.get(key)
        if ws is None:
            ws = torch.empty(need, dtype=torch.uint8, device=counts.device)
            _scan_ws[key] = ws
    L.call("sph_scan_exclusive", L.ptr(counts), int(n), L.ptr(out), L.ptr(ws), int(ws.numel()),
           L.stream_handle(stream))


def scan_ws_bytes(n: int) -> int:
    return int(L.load().sph_scan_workspace_bytes(n))


def sort_rows_(a: CSR, stream=None) -> None:
    c = a.c()
    L.call("sph_csr_sort_rows", C.byref(c), L.ptr(a.col), L.stream_handle(stream))


def transpose(a: CSR, ncols: int, stream=None) -> CSR:
    """Transpose with exact allocation; rows of the result sorted."""
    dev = a.rowptr.device
    nnz = a.nnz if a.nnz is not None else int(a.rowptr[a.nrows].item())
    t_counts = torch.empty(max(ncols, 1), dtype=torch.int32, device=dev)
    t_slot = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    t_rowptr = torch.empty(ncols + 1, dtype=torch.int32, device=dev)
    t_col = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    ws = torch.empty(scan_ws_bytes(ncols), dtype=torch.uint8, device=dev)
    c = a.c()
    L.call("sph_csr_transpose", C.byref(c), int(ncols), L.ptr(t_counts), L.ptr(t_slot),
           L.ptr(t_rowptr), L.ptr(t_col), L.ptr(ws), int(ws.numel()), L.stream_handle(stream))
    return CSR(t_rowptr, t_col, ncols, nnz)


def pair_geometry(a: CSR, xq, xg, hq, support: float, thresh: float = 1e-7, stream=None):
    """COO int64 rows/cols, unit vectors and q of every pair
    (neighborSearch.py:216-227 for hq=None, :243-250 otherwise)."""
    dev = a.rowptr.device
    nnz = a.nnz if a.nnz is not None else int(a.rowptr[a.nrows].item())
    rows = torch.empty(nnz, dtype=torch.int64, device=dev)
    cols = torch.empty(nnz, dtype=torch.int64, device=dev)
    rhat = torch.empty((nnz, 2), dtype=torch.float32, device=dev)
    q = torch.empty(nnz, dtype=torch.float32, device=dev)
    c = a.c()
    L.call("sph_pair_geometry", C.byref(c), L.ptr(xq), L.ptr(xg), L.ptr(hq), float(support),
           float(thresh), L.ptr(rows), L.ptr(cols), L.ptr(rhat), L.ptr(q), L.stream_handle(stream))
    return rows, cols, rhat, q
