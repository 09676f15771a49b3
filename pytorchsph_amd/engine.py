"""Device-resident batched SPH engine on the C-ABI library.

`SceneBatch` holds many independent scenes of one scheme in flat device
arrays (scene s owns fluid particles [p_begin, p_begin + n_s) and boundary
particles [b_begin, b_begin + nb_s)).  One step advances every scene:
  delta-SPH  RK4, four right-hand sides (simulationBase.py:679-705), each
             = grid + fluid pairs + ghost/boundary pairs, mDBC boundary
             density, one fused fluid gather (deltaSPH.py:146-200);
  DFSPH      semi-implicit Euler (simulationBase.py:647-652) around one
             right-hand side: summation density with the Akinci boundary,
             the relaxed-Jacobi density solve, viscosity terms
             (dfsph.py:134-202).
At the start of every step the particle state is permuted into the cell
order of the previous step's last neighbour search (each scene keeps its
index range; `uid` records where every particle came from), so that a
particle's neighbours sit next to it in memory and every gather reads
coalesced, cache-local data.  All buffers are allocated up front with fixed
pair capacities, so a step launches kernels only and can be captured into
a hipGraph.  A single plain scene is nscenes == 1.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _lib as L
from .scenes import Scene
from .search import CSR, NeighborGrid, scan_ws_bytes

DSPH_TERMS_MOG = (L.T_CONT | L.T_CONT_B | L.T_MOG | L.T_PRESS | L.T_PRESS_B | L.T_VISC |
                  L.T_VISC_B | L.T_LAM | L.T_GRAV)
DSPH_TERMS_DD = (DSPH_TERMS_MOG & ~L.T_MOG) | L.T_DDSPH
DIV_MAX_HIST = 16


def eos_host(phys: dict, rho) -> np.ndarray:
    """Equation of state on the host (pressure.py:99-111): Tait (eos 0),
    linear delta-SPH (1) or compressible (2) of the normalised density."""
    r = np.asarray(rho, dtype=np.float64)
    rho0, c0 = phys["rho0"], phys["c0"]
    e = int(phys.get("eos", 0))
    if e == 0:
        return rho0 * c0 * c0 / phys["gamma"] * (r ** phys["gamma"] - 1.0)
    if e == 1:
        return c0 * c0 * (r * rho0 - rho0)
    return phys["kappa"] * (r * rho0 - rho0)


def walk_segments(p_begin, n, run: int = 256) -> np.ndarray:
    """Runs of at most `run` sorted slots of one scene each, covering every
    scene's slots [p_begin, p_begin + n) in order: the workgroups of
    sph_dsph_cellwalk.  [nseg, 2] int32 {first, end}."""
    parts = []
    for a, c in zip(p_begin, n):
        st = np.arange(int(a), int(a) + int(c), run, dtype=np.int64)
        parts.append(np.stack([st, np.minimum(st + run, int(a) + int(c))], 1))
    if not parts:
        return np.zeros((0, 2), np.int32)
    return np.ascontiguousarray(np.concatenate(parts).astype(np.int32))


def phys_records(scenes) -> np.ndarray:
    rec = np.zeros(len(scenes), dtype=L.PHYS_DTYPE)
    for s, sc in enumerate(scenes):
        p = sc.phys
        for k in L.PHYS_DTYPE.names:
            if k == "pad":
                continue
            if k in p:
                rec[s][k] = p[k]
    return rec


class _PairBuf:
    """Fixed-capacity CSR for nrows query rows."""

    def __init__(self, nrows: int, cap: int, dev):
        self.nrows = nrows
        self.counts = torch.zeros(max(nrows, 1), dtype=torch.int32, device=dev)
        self.rowptr = torch.zeros(nrows + 1, dtype=torch.int32, device=dev)
        self.col = torch.zeros(max(cap, 1), dtype=torch.int32, device=dev)
        self.cap = max(cap, 1)

    def csr(self) -> CSR:
        return CSR(self.rowptr, self.col, self.nrows)


class SceneBatch:
    """Many scenes of one scheme ("dsph" or "dfsph"), resident on one GPU."""

    def __init__(self, scenes: list[Scene], device="cuda", pair_cap: float = 40.0,
                 bdy_pair_cap: float = 40.0, dfsph_chunk: int = 8, reorder: bool = True,
                 carry: list | None = None, fuse_rk4: bool | None = None,
                 cellwalk: bool = True, cell_div: int = 1, stencil_lists: bool = True):
        """fuse_rk4: delta-SPH stages run the RK4 combine and the next
        stage's records inside the gather (sph_dsph_gather_rk4 or
        sph_dsph_cellwalk_rk4) instead of as separate kernels; the results
        are the same, bit for bit.  None (the default): on for cell-walk
        batches, where it saves the pack and combine launches of every stage
        (0.21 ms of the 4M-particle dam-break step), off for list batches,
        where the fused gather's extra registers (67 instead of 60, seven
        wavefronts per SIMD instead of eight) cost what the launches save.

        cellwalk: delta-SPH stages with the default terms (MOG diffusion, no
        delta-plus) find their fluid pairs inside the gather
        (sph_dsph_cellwalk) and build no fluid pair list; the results are
        those of the list gather, bit for bit.  The boundary lists are built
        either way.

        cell_div: grid cells of support / cell_div searched with a stencil
        of radius cell_div (1 or 2).  The pair sets are the same; the order
        inside a row, and so the rounding of the sums, follows the stencil.
        The cell walk needs cell_div 1.

        stencil_lists: DFSPH steps permute the state into the order of the
        step's own grid and keep each particle's fluid neighbours as
        bitmasks over its stencil (sph_stencil_list_build) instead of a CSR
        list; the density, alpha / source, solve and finish gathers stage
        their rows' stencil records in LDS (the sph_*_sl entry points, one
        lane per row).  The results are those of the CSR gathers with one
        lane per row, bit for bit.  Needs reorder and cell_div 1.

        carry[k], when given, continues scene k from another batch: a dict
        from `export` with device tensors of its positions, velocities and
        normalised density, its particle ids and the cell order of that
        batch's last search.  The scene then advances exactly as it would
        have in the batch it came from (scene k's own x / v are not used)."""
        assert scenes, "empty batch"
        scheme = scenes[0].scheme
        assert all(s.scheme == scheme for s in scenes), "one scheme per batch"
        self.scheme = scheme
        self.scenes = scenes
        self.dev = dev = torch.device(device)
        self.nscenes = len(scenes)
        self.n = np.array([s.n for s in scenes])
        self.nb = np.array([s.nb for s in scenes])
        self.p_begin = np.concatenate([[0], np.cumsum(self.n)])[:-1]
        self.b_begin = np.concatenate([[0], np.cumsum(self.nb)])[:-1]
        NP, NB = int(self.n.sum()), int(self.nb.sum())
        self.NP, self.NB = NP, NB
        self.solid = scenes[0].boundary == "solid"
        self.dfsph_chunk = dfsph_chunk
        self.reorder = reorder
        self._sorted_valid = False
        f32 = torch.float32

        def cat(arrs):
            if not arrs or sum(len(a) for a in arrs) == 0:
                return torch.zeros(1, dtype=f32, device=dev)
            return torch.from_numpy(np.ascontiguousarray(np.concatenate(arrs))).to(dev)

        carry = carry if carry is not None else [None] * self.nscenes
        assert len(carry) == self.nscenes

        def state(key):
            if all(c is None for c in carry):
                return cat([getattr(s, key) for s in scenes]).float().contiguous()
            parts = [c[key].to(dev) if c is not None
                     else torch.from_numpy(np.ascontiguousarray(getattr(s, key), np.float32)).to(dev)
                     for s, c in zip(scenes, carry)]
            return torch.cat(parts).contiguous()

        # ---- fluid state (perennial)
        self.x = state("x")
        self.v = state("v")
        self.rho = torch.ones(NP, dtype=f32, device=dev)
        self.area = cat([np.full(s.n, s.area, np.float32) for s in scenes])
        self.rest = cat([np.full(s.n, s.rho0, np.float32) for s in scenes])
        self.h = cat([np.full(s.n, s.support, np.float32) for s in scenes])
        self.pscene = torch.from_numpy(np.repeat(np.arange(self.nscenes, dtype=np.int32),
                                                 self.n)).to(dev)
        self.uid = torch.arange(NP, dtype=torch.int64, device=dev)
        self.carry = carry
        # x, v, rho, uid blocks of sph_permute_fields (n * 7 words + 4 per field)
        self.tmp = torch.empty(7 * NP + 16, dtype=f32, device=dev)
        # ---- boundary state (static)
        self.bx = cat([s.bx for s in scenes]).float().contiguous()
        self.bgx = cat([s.bgx for s in scenes]).float().contiguous() if self.solid else None
        self.bvol = cat([s.bvol for s in scenes])
        self.brest = cat([np.full(s.nb, s.rho0, np.float32) for s in scenes])
        self.bh = cat([np.full(s.nb, s.support, np.float32) for s in scenes])
        self.bterm = None if self.solid else cat([s.bterm for s in scenes])
        self.bscene = torch.from_numpy(np.repeat(np.arange(self.nscenes, dtype=np.int32),
                                                 self.nb)).to(dev)
        self.brho = torch.ones(max(NB, 1), dtype=f32, device=dev)
        self.bp = torch.zeros(max(NB, 1), dtype=f32, device=dev)
        self.bcount = torch.zeros(max(NB, 1), dtype=torch.int32, device=dev)
        # ---- per-scene constants
        self.phys_host = phys_records(scenes)
        self.phys = L.records_to_tensor(self.phys_host, dev)
        d0 = scenes[0].phys
        # delta-plus: surface detection in every right-hand side with the
        # delta-plus pressure switch (scheme 'deltaPlus'), and the shift
        # after each step (shifting.enabled); one setting per batch
        self.dplus = bool(d0.get("dplus", False))
        self.shifting = bool(d0.get("shifting", False))
        assert all(bool(s.phys.get("dplus", False)) == self.dplus and
                   bool(s.phys.get("shifting", False)) == self.shifting for s in scenes), \
            "one delta-plus setting per batch"
        # DFSPH divergence solve before the density solve (divergenceSolver,
        # dfsph.py:175-179); one setting per batch
        self.divergence = scheme == "dfsph" and bool(d0.get("divergence", False))
        assert all(bool(s.phys.get("divergence", False)) == self.divergence
                   for s in scenes if s.scheme == "dfsph"), "one divergence setting per batch"
        if scheme == "dsph":
            self.terms = DSPH_TERMS_DD if d0.get("diffusion") == "deltaSPH" else DSPH_TERMS_MOG
            if not d0.get("boundary_diffusion", True):
                self.terms &= ~L.T_VISC_B

        assert cell_div in (1, 2), "cell_div is 1 or 2"
        self.cell_div = int(cell_div)
        self.cellwalk = (cellwalk and cell_div == 1 and scheme == "dsph"
                         and not (self.dplus or self.shifting) and not self.terms & L.T_DDSPH)
        self.fuse_rk4 = self.cellwalk if fuse_rk4 is None else bool(fuse_rk4)
        self.stencil_lists = (stencil_lists and scheme == "dfsph" and reorder and cell_div == 1)
        if self.cellwalk:
            # the cell walk's pair predicate takes the support from the
            # physics records; it must be the support the grid was sized for
            assert all(float(np.float32(ph["h"])) == float(np.float32(s.support))
                       for ph, s in zip(self.phys_host, scenes))

        # ---- neighbour search
        # cells of support / d, the table keeping the margins of the
        # reference's grid (qmin = min - h, qmax = max + 2 h)
        d = self.cell_div
        grid_scenes = [dict(vmin=np.asarray(s.vmin, np.float64) - s.support * (1 - 1 / d),
                            vmax=np.asarray(s.vmax, np.float64) + 2 * s.support * (1 - 1 / d),
                            hmax=s.support / d, p_begin=int(pb), p_count=int(s.n))
                       for s, pb in zip(scenes, self.p_begin)]
        self.grid = NeighborGrid(grid_scenes, NP, dev)
        # delta-SPH batches: the static boundary points sorted once into a
        # grid of the same cells; each stage finds its fluid -> boundary
        # pairs in it (in the cell walk, or for the list by a search with
        # the fluid points as queries), so list and cell walk add them in
        # the same order.  The boundary -> fluid list is built only where
        # it is read: DFSPH's Akinci rows and delta-plus's wall rows.
        self.bgrid = None
        self.needs_bf = scheme == "dfsph" or self.dplus or self.shifting
        if scheme == "dsph" and self.solid and NB:
            bscenes = [dict(gs, p_begin=int(bb), p_count=int(s.nb))
                       for gs, bb, s in zip(grid_scenes, self.b_begin, scenes)]
            self.bgrid = NeighborGrid(bscenes, NB, dev)
            if dev.type == "cuda":      # not on the meta device of footprint()
                self.bgrid.build(self.bx, self.bscene)
        self.ff = _PairBuf(NP, int(math.ceil(pair_cap * NP)), dev)
        bcap = int(math.ceil(bdy_pair_cap * max(NB, 1)))
        self.bf = _PairBuf(NB, bcap, dev)
        self.gf = _PairBuf(NB, bcap, dev) if self.solid else None
        self.fb = _PairBuf(NP, bcap, dev)
        self.t_slot = torch.zeros(bcap, dtype=torch.int32, device=dev)
        if self.stencil_lists:
            self.sl_runs = torch.zeros((max(NP, 1), 4), dtype=torch.int32, device=dev)
            self.sl_masks = torch.zeros((3, max(NP, 1)), dtype=torch.int64, device=dev)
        segs = walk_segments(self.p_begin, self.n)
        self.nseg = len(segs)
        self.segs = torch.from_numpy(segs).to(dev) if self.cellwalk and self.nseg else None
        self.scan_ws = torch.zeros(scan_ws_bytes(max(NP, NB, 1)), dtype=torch.uint8, device=dev)
        # ---- step buffers
        self.acc = torch.zeros((NP, 2), dtype=f32, device=dev)
        self.dpdt = torch.zeros(NP, dtype=f32, device=dev)
        self.frec = torch.zeros((max(NP, 1), 8), dtype=f32, device=dev)
        self.brec = torch.zeros((max(NB, 1), 8), dtype=f32, device=dev)
        if scheme == "dsph":
            self.S = [(torch.empty_like(self.x), torch.empty_like(self.v), torch.empty_like(self.rho))
                      for _ in range(2)]
            self.K = torch.zeros((NP, 5), dtype=f32, device=dev)
            # records of the next RK4 stage while the gather reads this one's
            self.frec2 = torch.zeros_like(self.frec) if self.fuse_rk4 else None
            self.gradrho = torch.zeros((NP, 2), dtype=f32, device=dev)
            self.Lmat = torch.zeros((NP, 4), dtype=f32, device=dev)
            if self.dplus or self.shifting:
                assert self.solid and not self.fuse_rk4, \
                    "delta-plus runs with mDBC walls and the separate RK4 kernels"
                self._alloc_dplus()
        else:
            self.vol = torch.zeros(NP, dtype=f32, device=dev)
            self._alloc_dfsph()
        self.steps_done = 0
        self.graph = None
        if any(c is not None for c in self.carry):
            self._apply_carry()

    def _apply_carry(self):
        """Density, particle ids, pending cell order and status bits of
        carried scenes; fresh scenes get the identity order, so the first
        reorder leaves them as they are.  The status bits a scene raised in
        earlier batches stay raised (device status words are sticky)."""
        order = torch.arange(self.NP, dtype=torch.int32, device=self.dev)
        grid_bits = L.ST_OUTSIDE | L.ST_OVERFLOW
        for k, c in enumerate(self.carry):
            if c is None:
                continue
            a, n = int(self.p_begin[k]), int(self.n[k])
            self.rho[a:a + n].copy_(c["rho"])
            self.uid[a:a + n].copy_(c["uid"] + a)
            order[a:a + n].copy_(c["order"] + a)
            st = int(c.get("status", 0))
            if st & grid_bits:
                self.grid.status[k] |= st & grid_bits
            if self.scheme == "dfsph" and st & ~grid_bits:
                self.df["status"][k] |= st & ~grid_bits
        self.grid.sorted_idx[: self.NP].copy_(order)
        self._sorted_valid = True

    @staticmethod
    def footprint(scenes: list[Scene], **kw) -> int:
        """Device bytes a batch of these scenes allocates (built on the meta
        device: no memory is touched)."""
        b = SceneBatch(scenes, device="meta", **kw)
        seen, total = set(), 0

        def walk(o, depth):
            nonlocal total
            if isinstance(o, torch.Tensor):
                if id(o) not in seen:
                    seen.add(id(o))
                    total += o.numel() * o.element_size()
            elif isinstance(o, dict):
                for v in o.values():
                    walk(v, depth + 1)
            elif isinstance(o, (list, tuple)):
                for v in o:
                    walk(v, depth + 1)
            elif depth < 3 and hasattr(o, "__dict__") and not isinstance(o, type):
                for v in vars(o).values():
                    walk(v, depth + 1)

        walk(b, 0)
        return total

    def export(self, k: int) -> tuple[Scene, dict]:
        """Scene k with the carry record (device copies of its state in the
        current particle order) that lets another batch continue it
        exactly."""
        a, n = int(self.p_begin[k]), int(self.n[k])
        order = (self.grid.sorted_idx[a:a + n] - a if self._sorted_valid
                 else torch.arange(n, dtype=torch.int32, device=self.dev))
        carry = dict(x=self.x[a:a + n].clone(), v=self.v[a:a + n].clone(),
                     rho=self.rho[a:a + n].clone(), uid=self.uid[a:a + n] - a,
                     order=order.clone(), status=self.scene_status(k))
        return self.scenes[k], carry

    def scene_status(self, k: int) -> int:
        st = int(self.grid.status[k].item())
        if self.bgrid is not None:
            st |= int(self.bgrid.status[k].item())
        if self.scheme == "dfsph":
            st |= int(self.df["status"][k].item())
        return st

    # ----------------------------------------------------------- helpers
    def _alloc_dplus(self):
        """Buffers of the delta-plus passes, the static boundary->boundary
        pair list (solidBoundary.py:180) and the map of quirk deltaPlus.py:151
        (boundary neighbour b takes the volume of the scene's fluid particle
        with original index b)."""
        from .scenes import lattice_calibration

        dev, NP, NB, f32 = self.dev, self.NP, self.NB, torch.float32
        nb1 = max(NB, 1)
        self.eig = torch.zeros((NP, 2), dtype=f32, device=dev)
        self.normal = torch.zeros((NP, 2), dtype=f32, device=dev)
        self.smask = torch.full((NP,), 2.0, dtype=f32, device=dev)
        self.mask_tmp = torch.zeros(NP, dtype=f32, device=dev)
        self.vf = torch.zeros(NP, dtype=f32, device=dev)
        self.vb = torch.zeros(nb1, dtype=f32, device=dev)
        self.bvol_n = torch.zeros(nb1, dtype=f32, device=dev)
        self.beig = torch.zeros((nb1, 2), dtype=f32, device=dev)
        self.bL = torch.zeros((nb1, 4), dtype=f32, device=dev)
        self.shift = torch.zeros((NP, 2), dtype=f32, device=dev)
        self.vmax_ws = torch.zeros(self.nscenes, dtype=torch.int32, device=dev)
        self.slot_of = torch.zeros(NP, dtype=torch.int64, device=dev)
        self.arange_np = torch.arange(NP, dtype=torch.int64, device=dev)
        calib = lattice_calibration()[0]
        self.packing = torch.tensor([math.sqrt(s.area / calib) / s.support for s in self.scenes],
                                    dtype=f32, device=dev)
        # original fluid index (global) of boundary point b's namesake, or -1
        orig = np.full(nb1, -1, dtype=np.int64)
        for k, s in enumerate(self.scenes):
            m = min(int(s.n), int(s.nb))
            b0, p0 = int(self.b_begin[k]), int(self.p_begin[k])
            orig[b0:b0 + m] = p0 + np.arange(m)
        self.b_orig = torch.from_numpy(orig).to(dev)
        self.b_valid = self.b_orig >= 0
        self.b_orig_c = self.b_orig.clamp(min=0)
        if dev.type == "meta" or NB == 0:
            cap = int(40 * nb1)
            self.bb = CSR(torch.zeros(nb1 + 1, dtype=torch.int32, device=dev),
                          torch.zeros(cap, dtype=torch.int32, device=dev), NB)
            return
        gs = [dict(vmin=s.vmin, vmax=s.vmax, hmax=s.support, p_begin=int(b0), p_count=int(s.nb))
              for s, b0 in zip(self.scenes, self.b_begin)]
        g = NeighborGrid(gs, NB, dev)
        g.build(self.bx, self.bscene)
        self.bb = g.search(self.bx, self.bh, self.bscene, 1, self_sorted=True, sort_rows=True)
        del g

    def _dplus_detect(self, F, ff, fbp, x, rho, sh):
        """detectSurface of the stage state (deltaPlus.py:242-245): lambda of
        fluid and boundary rows, normals, surface classes."""
        B = self.boundary()
        L.call("sph_dsph_normalization", C.byref(F), L.ptr(self.phys), C.byref(ff), C.byref(B),
               fbp, None, None, L.ptr(self.Lmat), L.ptr(self.eig), sh)
        torch.div(self.area, rho, out=self.vf)
        bx, beig, bvol = None, None, None
        if self.NB:
            torch.div(self.bvol, self.brho[: self.NB], out=self.vb[: self.NB])
            bf = self.bf.csr().c()
            bb = self.bb.c()
            L.call("sph_normalization_rows", L.ptr(self.bx), L.ptr(self.bscene), self.NB,
                   L.ptr(self.phys), C.byref(bf), L.ptr(x), L.ptr(self.vf), C.byref(bb),
                   L.ptr(self.bx), L.ptr(self.vb), None, L.ptr(self.bL), L.ptr(self.beig), sh)
            # deltaPlus.py:151: fluid arrays indexed by the boundary index,
            # through the particle ids of the cell-ordered state
            self.slot_of.index_copy_(0, self.uid, self.arange_np)
            v = self.vf[self.slot_of[self.b_orig_c]]
            torch.where(self.b_valid, v, torch.zeros_like(v), out=self.bvol_n)
            bx, beig, bvol = self.bx, self.beig, self.bvol_n
        L.call("sph_dplus_detect", C.byref(F), L.ptr(self.phys), C.byref(ff), L.ptr(self.Lmat),
               L.ptr(self.eig), L.ptr(bx), L.ptr(beig), L.ptr(bvol), fbp, None,
               L.ptr(self.normal), L.ptr(self.mask_tmp), None, None, L.ptr(self.smask), sh)

    def _dplus_shift(self, s=None):
        """Step 14 of integrate (simulationBase.py:708-732): search and mDBC
        at the integrated positions, then the shift with the last stage's
        normals and classes, added to the positions."""
        sh = L.stream_handle(s)
        self.search(self.x, s)
        F = self.fluid(self.x, self.v, self.rho)
        ff = self.ff.csr().c()
        fb = self.fb.csr().c() if self.NB else None
        if self.NB:
            gf = self.gf.csr().c()
            L.call("sph_mdbc_density", C.byref(F), L.ptr(self.phys), L.ptr(self.bx),
                   L.ptr(self.bgx), L.ptr(self.bscene), self.NB, C.byref(gf), L.ptr(self.brho),
                   L.ptr(self.bp), L.ptr(self.bcount), sh)
        B = self.boundary()
        L.call("sph_dplus_shift", C.byref(F), L.ptr(self.phys), self.nscenes, L.ptr(self.packing),
               C.byref(ff), C.byref(B), C.byref(fb) if fb is not None else None,
               L.ptr(self.normal), L.ptr(self.smask), L.ptr(self.vmax_ws), L.ptr(self.shift), sh)
        self.x.add_(self.shift)

    def _alloc_dfsph(self):
        dev, NP, NB, S = self.dev, self.NP, self.NB, self.nscenes
        z = lambda *s, dt=torch.float32: torch.zeros(*s, dtype=dt, device=dev)  # noqa: E731
        nb1 = max(NB, 1)
        self.max_hist = 260
        ints = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)  # noqa
        self.df = dict(
            vpred=z(NP, 2), aact=z(NP), alpha=z(NP), src=z(NP), p=z(NP), p2=z(NP), apred=z(NP, 2),
            resid=z(NP), xrec=z(NP, 4), arec=z(NP, 4), prec=z(NP, 4), wvec=z(NP, 2), b_aact=z(nb1), b_alpha=z(nb1),
            b_src=z(nb1), b_p=z(nb1), b_resid=z(nb1), bxrec=z(nb1, 4),
            err_acc=z(L.DFSPH_ERR_SLOTS, L.dfsph_err_stride(S), dt=torch.float64), err_hist=z(S, self.max_hist, dt=torch.float64),
            iters=z(S, dt=torch.int32), active=z(S, dt=torch.int32),
            any_active=z(1, dt=torch.int32), status=z(S, dt=torch.int32),
            nfluid=ints(self.n), nbdy=ints(self.nb), fbegin=ints(self.p_begin),
            bbegin=ints(self.b_begin), act_scene=z(S, dt=torch.int32),
            act_pre=z(2 * (S + 1), dt=torch.int32), act_n=z(3, dt=torch.int32))
        self.any_active_host = torch.zeros(1, dtype=torch.int32,
                                           pin_memory=self.dev.type == "cuda")
        self.iters_host = torch.zeros(S, dtype=torch.int32, pin_memory=self.dev.type == "cuda")
        # iterations the next step's first chunk runs: the last step's most
        # iterated scene (at least the stopping rule's minimum); the loop
        # continues in chunks of dfsph_chunk when a scene needs more
        self.min_iters = int(max(1, min(self.phys_host["min_iters"])))
        self.pred_iters = self.min_iters
        if self.divergence:
            # the divergence solve's own iteration counters and residual
            # history, and the limits it runs with (divergenceThreshold,
            # min/maxDivergenceSolverIterations, dfsph.py:240)
            self.div_iters = z(S, dt=torch.int32)
            self.div_err_hist = z(S, DIV_MAX_HIST, dt=torch.float64)
            ph = self.phys_host.copy()
            for k, sc in enumerate(self.scenes):
                ph["tau"][k] = sc.phys.get("div_tau", 1e-2)
                ph["min_iters"][k] = sc.phys.get("div_min_iters", 2)
                ph["max_iters"][k] = sc.phys.get("div_max_iters", 8)
            self.phys_div = L.records_to_tensor(ph, dev)
            self.div_chunk = int(max(ph["max_iters"])) + 1

    def _dfsph_struct(self, divergence: bool = False) -> L.SphDfsph:
        ptrs = {k: L.ptr(self.df[k]) for k in L.DFSPH_FIELDS}
        hist = self.max_hist
        if divergence:
            ptrs["iters"] = L.ptr(self.div_iters)
            ptrs["err_hist"] = L.ptr(self.div_err_hist)
            hist = DIV_MAX_HIST
        return L.SphDfsph(*[ptrs[k] for k in L.DFSPH_FIELDS], self.nscenes, hist,
                          L.dfsph_grid_blocks(self.NP + self.NB), 1 if divergence else 0)

    def fluid(self, x, v, rho, vol=None) -> L.SphFluid:
        return L.SphFluid(L.ptr(x), L.ptr(v), L.ptr(rho), L.ptr(self.area), L.ptr(self.rest),
                          L.ptr(vol), L.ptr(self.pscene), self.NP)

    def boundary(self) -> L.SphBoundary:
        return L.SphBoundary(L.ptr(self.bx), None, L.ptr(self.brho), L.ptr(self.bvol),
                             L.ptr(self.bp), L.ptr(self.brest), L.ptr(self.bscene), self.NB)

    def _pairs(self, buf: _PairBuf, q: L.SphQuery, s):
        g = self.grid
        g.count(q, buf.counts, s)
        L.call("sph_scan_exclusive", L.ptr(buf.counts), buf.nrows, L.ptr(buf.rowptr),
               L.ptr(self.scan_ws), int(self.scan_ws.numel()), L.stream_handle(s))
        g.fill(q, buf.rowptr, buf.col, s)

    def search(self, x, s=None, fluid_pairs: bool = True, ghost_pairs: bool = True,
               bdy_pairs: bool = True, build: bool = True):
        """Grid + every pair list of the current positions (without the
        fluid list when not fluid_pairs, the ghost list when not
        ghost_pairs, the boundary lists when not bdy_pairs).  Delta-SPH
        batches build the fluid -> boundary list by a search of the fluid
        points in the static boundary grid, so its rows follow that grid's
        order (the order the cell walk walks them in); DFSPH batches
        transpose the boundary -> fluid list (rows by boundary index).
        build=False: the grid is already built (and its points sorted)."""
        g = self.grid
        if build:
            g.build(x, self.pscene, s)
        if fluid_pairs:
            self._pairs(self.ff, g.query_struct(x, self.h, self.pscene, self.cell_div, True), s)
        if self.NB:
            if self.solid and ghost_pairs:
                self._pairs(self.gf, g.query_struct(self.bgx, self.bh, self.bscene, self.cell_div,
                                                    False), s)
            if bdy_pairs and self.needs_bf:
                self._pairs(self.bf, g.query_struct(self.bx, self.bh, self.bscene,
                                                    self.cell_div, False), s)
            if bdy_pairs and self.bgrid is not None:
                bg = self.bgrid
                q = bg.query_struct(x, self.h, self.pscene, self.cell_div, False)
                bg.count(q, self.fb.counts, s)
                L.call("sph_scan_exclusive", L.ptr(self.fb.counts), self.NP, L.ptr(self.fb.rowptr),
                       L.ptr(self.scan_ws), int(self.scan_ws.numel()), L.stream_handle(s))
                bg.fill(q, self.fb.rowptr, self.fb.col, s)
            elif bdy_pairs:
                c = self.bf.csr().c()
                L.call("sph_csr_transpose", C.byref(c), self.NP, L.ptr(self.fb.counts),
                       L.ptr(self.t_slot), L.ptr(self.fb.rowptr), L.ptr(self.fb.col),
                       L.ptr(self.scan_ws), int(self.scan_ws.numel()), L.stream_handle(s))
        self._sorted_valid = True

    def _bgrid_arg(self):
        """The static boundary grid as the cell walk takes it (None: no
        boundary points)."""
        if self.bgrid is None:
            return None
        self._bgrid_c = self.bgrid._grid()
        return C.byref(self._bgrid_c)

    def _order(self):
        # with the state in cell order, index order is already local
        return None if self.reorder else L.ptr(self.grid.sorted_idx)

    def reorder_state(self, s=None):
        """Permutes the perennial particle state into the cell order of the
        last neighbour search.  The permutation keeps every scene's range, so
        the per-scene constants (area, rest density, support) stay put; x, v,
        rho and the ids move together in two launches."""
        if not (self.reorder and self._sorted_valid and self.NP):
            return
        fl = L.SphFields()
        for q, (t, w) in enumerate(((self.x, 2), (self.v, 2), (self.rho, 1), (self.uid, 2))):
            fl.ptr[q] = L.ptr(t)
            fl.width[q] = w
        fl.nfields = 4
        L.call("sph_permute_fields", C.byref(fl), L.ptr(self.grid.sorted_idx), self.NP,
               L.ptr(self.tmp), L.stream_handle(s))

    # ---------------------------------------------------------- delta-SPH
    def _dsph_prepare(self, x, v, rho, frec, s=None):
        """Search, mDBC boundary density, delta-SPH density gradient and the
        records (fluid records into frec unless it is None) of the state
        (x, v, rho).  Returns the gather's arguments."""
        # cell-walk batches find every pair in the kernels that use it: the
        # stage builds the fluid grid and no list
        w = self.cellwalk
        self.search(x, s, fluid_pairs=not w, ghost_pairs=not w, bdy_pairs=not w)
        F = self.fluid(x, v, rho)
        sh = L.stream_handle(s)
        ff = self.ff.csr().c()
        fb = self.fb.csr().c() if self.NB else None
        if self.NB and self.cellwalk:
            # the walk writes the boundary records too (walls at rest)
            g = self.grid._grid()
            L.call("sph_mdbc_density_walk_records", C.byref(F), L.ptr(self.phys), L.ptr(self.bx),
                   None, L.ptr(self.bvol), L.ptr(self.bgx), L.ptr(self.bh), L.ptr(self.bscene),
                   self.NB, C.byref(g), L.ptr(self.brho), L.ptr(self.bp), L.ptr(self.bcount),
                   L.ptr(self.brec), sh)
        elif self.NB:
            gf = self.gf.csr().c()
            L.call("sph_mdbc_density", C.byref(F), L.ptr(self.phys), L.ptr(self.bx),
                   L.ptr(self.bgx), L.ptr(self.bscene), self.NB, C.byref(gf), L.ptr(self.brho),
                   L.ptr(self.bp), L.ptr(self.bcount), sh)
        B = self.boundary()
        fbp = C.byref(fb) if fb is not None else None
        if self.dplus or self.shifting:
            self._dplus_detect(F, ff, fbp, x, rho, sh)
        if self.terms & L.T_DDSPH:
            L.call("sph_dsph_normalization", C.byref(F), L.ptr(self.phys), C.byref(ff), C.byref(B),
                   fbp, self._order(), None, L.ptr(self.Lmat), None, sh)
            L.call("sph_dsph_density_gradient", C.byref(F), L.ptr(self.phys), C.byref(ff),
                   C.byref(B), fbp, L.ptr(self.Lmat), self._order(), L.ptr(self.gradrho), sh)
        # boundary records unless the mDBC walk wrote them
        packed_b = bool(self.NB and self.cellwalk)
        L.call("sph_dsph_pack", C.byref(F), L.ptr(self.phys), None if packed_b else C.byref(B),
               L.ptr(frec) if frec is not None else None, L.ptr(self.brec), sh)
        return ff, fbp

    def dsph_rhs(self, x, v, rho, s=None):
        """dpdt / acc of the state (x, v, rho) into self.dpdt / self.acc."""
        ff, fbp = self._dsph_prepare(x, v, rho, self.frec, s)
        grr = L.ptr(self.gradrho) if self.terms & L.T_DDSPH else None
        if self.dplus:
            # the pressure terms with the delta-plus switch
            # (computePressureAccelDeltaPlus, pressure.py:40-52)
            L.call("sph_dsph_gather_packed_dplus", L.ptr(self.frec), L.ptr(self.brec),
                   L.ptr(self.pscene), self.NP, L.ptr(self.phys), C.byref(ff), fbp, grr,
                   self.terms, self._order(), L.ptr(self.smask), L.ptr(self.dpdt),
                   L.ptr(self.acc), L.stream_handle(s))
            return
        if self.cellwalk:
            g = self.grid._grid()
            L.call("sph_dsph_cellwalk", C.byref(g), L.ptr(self.segs), self.nseg,
                   L.ptr(self.frec), L.ptr(self.brec), L.ptr(self.phys), None, self._bgrid_arg(),
                   self.terms, L.ptr(self.dpdt), L.ptr(self.acc), L.stream_handle(s))
            return
        L.call("sph_dsph_gather_packed", L.ptr(self.frec), L.ptr(self.brec), L.ptr(self.pscene),
               self.NP, L.ptr(self.phys), C.byref(ff), fbp, grr, self.terms, self._order(),
               L.ptr(self.dpdt), L.ptr(self.acc), L.stream_handle(s))

    def _dsph_step(self, s=None):
        self.reorder_state(s)
        P = (self.x, self.v, self.rho)
        S0, S1 = self.S
        stages = [(P, S0), (S0, S1), (S1, S0), (S0, P)]
        sh = L.stream_handle(s)
        if self.fuse_rk4:
            # stage k gathers recs[(k - 1) % 2] and writes the next stage's
            # records into recs[k % 2]; stage 1 packs its own
            recs = (self.frec, self.frec2)
            for k, (cur, out) in enumerate(stages, start=1):
                frec, nxt = recs[(k - 1) % 2], recs[k % 2]
                ff, fbp = self._dsph_prepare(*cur, frec if k == 1 else None, s)
                # stages 1-3 need no velocity array: the next stage reads the
                # velocities from the records
                rk = L.SphRk4Stage(k, L.ptr(self.x), L.ptr(self.v), L.ptr(self.rho),
                                   L.ptr(self.area), L.ptr(self.K), L.ptr(out[0]),
                                   L.ptr(out[1]) if k == 4 else None,
                                   L.ptr(out[2]), L.ptr(self.acc) if k == 4 else None,
                                   L.ptr(nxt) if k < 4 else None)
                if self.cellwalk:
                    g = self.grid._grid()
                    L.call("sph_dsph_cellwalk_rk4", C.byref(g), L.ptr(self.segs), self.nseg,
                           L.ptr(frec), L.ptr(self.brec), L.ptr(self.phys), None,
                           self._bgrid_arg(), self.terms, C.byref(rk), sh)
                    continue
                L.call("sph_dsph_gather_rk4", L.ptr(frec), L.ptr(self.brec), L.ptr(self.pscene),
                       self.NP, L.ptr(self.phys), C.byref(ff), fbp,
                       L.ptr(self.gradrho) if self.terms & L.T_DDSPH else None, self.terms,
                       self._order(), C.byref(rk), sh)
            return
        for k, (cur, out) in enumerate(stages, start=1):
            self.dsph_rhs(*cur, s=s)
            L.call("sph_integrate_rk4", k, L.ptr(self.pscene), L.ptr(self.phys), self.NP,
                   L.ptr(self.x), L.ptr(self.v), L.ptr(self.rho), L.ptr(cur[1]), L.ptr(self.acc),
                   L.ptr(self.dpdt), L.ptr(self.K), L.ptr(out[0]), L.ptr(out[1]), L.ptr(out[2]),
                   L.ptr(self.acc) if k == 4 else None, sh)
        if self.shifting:
            self._dplus_shift(s)

    # -------------------------------------------------------------- DFSPH
    def stencil_list(self) -> L.SphStencilList:
        return L.SphStencilList(L.ptr(self.sl_runs), L.ptr(self.sl_masks), self.NP)

    def _dfsph_search(self, s=None):
        """The step's grid and pair lists.  Stencil-list batches build the
        grid first, permute the state into its order (so a point's index is
        its sorted slot) and then list each point's fluid neighbours as
        stencil masks; the others permute into the previous search's order
        and build the CSR fluid list."""
        if not self.stencil_lists:
            self.reorder_state(s)
            self.search(self.x, s)
            return
        sh = L.stream_handle(s)
        g = self.grid
        g.build(self.x, self.pscene, s)
        self._sorted_valid = True
        self.reorder_state(s)
        L.call("sph_iota", L.ptr(g.sorted_idx), self.NP, sh)
        gc = g._grid()
        L.call("sph_stencil_list_build", C.byref(gc), L.ptr(self.h), L.ptr(self.sl_runs),
               L.ptr(self.sl_masks), sh)
        self.search(self.x, s, fluid_pairs=False, build=False)

    def _dfsph_pre(self, s=None):
        sh = L.stream_handle(s)
        self._dfsph_search(s)
        F = self.fluid(self.x, self.v, self.rho)
        ff = self.ff.csr().c()
        fb = self.fb.csr().c() if self.NB else None
        B = self.boundary()
        fbp = C.byref(fb) if fb is not None else None
        if self.stencil_lists:
            sl = self.stencil_list()
            L.call("sph_fluid_density_sl", C.byref(F), L.ptr(self.phys), C.byref(sl), C.byref(B),
                   fbp, L.ptr(self.rho), L.ptr(self.vol), sh)
        else:
            L.call("sph_fluid_density", C.byref(F), L.ptr(self.phys), C.byref(ff), C.byref(B),
                   fbp, self._order(), L.ptr(self.rho), L.ptr(self.vol), sh)
        if self.NB:
            bf = self.bf.csr().c()
            L.call("sph_akinci_boundary_density", C.byref(F), L.ptr(self.phys), L.ptr(self.bx),
                   L.ptr(self.bscene), self.NB, L.ptr(self.bterm), C.byref(bf), L.ptr(self.brho), sh)
        L.call("sph_dsph_gather", C.byref(F), L.ptr(self.phys), C.byref(ff), None, None, None,
               L.T_GRAV, None, None, L.ptr(self.acc), sh)
        if self.divergence:
            self._divergence_solve(F, ff, sh)
        D = self._dfsph_struct()
        L.call("sph_dfsph_prepare", C.byref(F), L.ptr(self.acc), L.ptr(self.phys), C.byref(B),
               C.byref(D), sh)
        bfc = self.bf.csr().c()
        if self.stencil_lists:
            L.call("sph_dfsph_alpha_source_sl", C.byref(F), L.ptr(self.phys), C.byref(sl),
                   C.byref(B), fbp, C.byref(bfc), C.byref(D), sh)
        else:
            L.call("sph_dfsph_alpha_source", C.byref(F), L.ptr(self.phys), C.byref(ff),
                   C.byref(B), fbp, C.byref(bfc), C.byref(D), sh)

    def _divergence_solve(self, F, ff, sh):
        """divergenceSolver (dfsph.py:175-179, 236-260): fluid-only solve on
        v* = v + dt a, then acc += its pressure acceleration.  At most
        maxDivergenceSolverIterations + 1 iterations, so one fixed chunk
        covers it and no host poll is needed; scenes that stop early are
        idle for the rest of the chunk."""
        D = self._dfsph_struct(divergence=True)
        pd = L.ptr(self.phys_div)
        L.call("sph_dfsph_prepare", C.byref(F), L.ptr(self.acc), pd, None, C.byref(D), sh)
        if self.stencil_lists:
            sl = self.stencil_list()
            L.call("sph_dfsph_alpha_source_sl", C.byref(F), pd, C.byref(sl), None, None, None,
                   C.byref(D), sh)
            for _ in range(self.div_chunk):
                L.call("sph_dfsph_iterate_sl", C.byref(F), pd, C.byref(sl), None, None, None,
                       C.byref(D), sh)
            L.call("sph_dfsph_finish_sl", C.byref(F), pd, C.byref(sl), None, C.byref(D),
                   L.ptr(self.acc), sh)
            return
        L.call("sph_dfsph_alpha_source", C.byref(F), pd, C.byref(ff), None, None, None,
               C.byref(D), sh)
        for _ in range(self.div_chunk):
            L.call("sph_dfsph_iterate", C.byref(F), pd, C.byref(ff), None, None, None,
                   C.byref(D), sh)
        L.call("sph_dfsph_finish", C.byref(F), pd, C.byref(ff), None, C.byref(D),
               L.ptr(self.acc), sh)

    def _dfsph_iters(self, k, s=None):
        sh = L.stream_handle(s)
        F = self.fluid(self.x, self.v, self.rho)
        ff = self.ff.csr().c()
        fb = self.fb.csr().c() if self.NB else None
        bfc = self.bf.csr().c()
        B = self.boundary()
        D = self._dfsph_struct()
        fbp = C.byref(fb) if fb is not None else None
        if self.stencil_lists:
            sl = self.stencil_list()
            for _ in range(k):
                L.call("sph_dfsph_iterate_sl", C.byref(F), L.ptr(self.phys), C.byref(sl),
                       C.byref(B), fbp, C.byref(bfc), C.byref(D), sh)
            return
        for _ in range(k):
            L.call("sph_dfsph_iterate", C.byref(F), L.ptr(self.phys), C.byref(ff), C.byref(B),
                   fbp, C.byref(bfc), C.byref(D), sh)

    def _dfsph_post(self, s=None):
        sh = L.stream_handle(s)
        ff = self.ff.csr().c()
        fb = self.fb.csr().c() if self.NB else None
        F = self.fluid(self.x, self.v, self.rho)
        D = self._dfsph_struct()
        # the final pressure acceleration, then velocity diffusion + laminar
        # with V = A / fluid-only density (density.py:53), in one walk of the
        # fluid list
        Fv = self.fluid(self.x, self.v, self.rho, self.vol)
        L.call("sph_dsph_pack", C.byref(Fv), L.ptr(self.phys), None, L.ptr(self.frec), None, sh)
        fbp = C.byref(fb) if fb is not None else None
        if self.stencil_lists:
            sl = self.stencil_list()
            L.call("sph_dfsph_finish_visc_sl", C.byref(F), L.ptr(self.phys), C.byref(sl), fbp,
                   C.byref(D), L.ptr(self.frec), L.T_VISC | L.T_LAM, L.ptr(self.acc), sh)
        else:
            L.call("sph_dfsph_finish_visc", C.byref(F), L.ptr(self.phys), C.byref(ff), fbp,
                   C.byref(D), L.ptr(self.frec), L.T_VISC | L.T_LAM, L.ptr(self.acc), sh)
        L.call("sph_integrate_semi_implicit", L.ptr(self.pscene), L.ptr(self.phys), self.NP,
               L.ptr(self.x), L.ptr(self.v), None, L.ptr(self.v), L.ptr(self.acc), None,
               L.ptr(self.x), L.ptr(self.v), None, sh)

    # iteration counts with a captured chunk graph: a step's first chunk runs
    # the smallest of these that covers the previous step's iterations
    CHUNK_SIZES = (1, 2, 3, 4, 6, 8, 12, 16)

    def _first_chunk(self) -> int:
        want = max(self.pred_iters, self.min_iters)
        fits = [k for k in self.CHUNK_SIZES if k >= want and k <= self.dfsph_chunk]
        return fits[0] if fits else self.dfsph_chunk

    def _dfsph_loop(self, s=None, graphs=None):
        """Iterates in chunks until every scene has converged: first the
        chunk that covers the last step's iterations, then chunks of
        dfsph_chunk (converged scenes cost nothing in either; a chunk longer
        than the solve costs its empty launches, a shorter one a host poll).
        The per-scene iteration counts come back with the poll."""
        stream = s if s is not None else torch.cuda.current_stream()
        k = self._first_chunk()
        while True:
            if graphs is not None:
                graphs[k].replay()
            else:
                self._dfsph_iters(k, s)
            self.any_active_host.copy_(self.df["any_active"], non_blocking=True)
            self.iters_host.copy_(self.df["iters"], non_blocking=True)
            stream.synchronize()
            if int(self.any_active_host[0]) == 0:
                break
            k = self.dfsph_chunk
        self.pred_iters = int(self.iters_host.max())

    # ----------------------------------------------------------- stepping
    def step(self, s=None):
        """One time step of every scene (eager launches)."""
        if self.scheme == "dsph":
            self._dsph_step(s)
        else:
            self._dfsph_pre(s)
            self._dfsph_loop(s)
            self._dfsph_post(s)
        self.steps_done += 1

    def capture(self, warm: bool = True):
        """Captures the step into hipGraphs (delta-SPH: one graph per step;
        DFSPH: prologue, a chunk of iterations, epilogue).  With `warm`, runs
        one eager step first (every scene advances by one step).  Without
        it nothing advances; a batch that has no cell order yet gets the
        identity order, so the recorded step reorders (a no-op the first
        time) exactly like every later one."""
        torch.cuda.synchronize()
        if warm:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self.step(s)
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
        elif not self._sorted_valid and self.NP:
            self.grid.sorted_idx[: self.NP].copy_(
                torch.arange(self.NP, dtype=torch.int32, device=self.dev))
            self._sorted_valid = True
            torch.cuda.synchronize()
        if self.scheme == "dsph":
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._dsph_step()
            self.graph = (g,)
        else:
            g1, g3 = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
            with torch.cuda.graph(g1):
                self._dfsph_pre()
            chunks = {}
            for k in sorted({*[c for c in self.CHUNK_SIZES if c <= self.dfsph_chunk],
                             self.dfsph_chunk}):
                chunks[k] = torch.cuda.CUDAGraph()
                with torch.cuda.graph(chunks[k]):
                    self._dfsph_iters(k)
            with torch.cuda.graph(g3):
                self._dfsph_post()
            self.graph = (g1, chunks, g3)
        torch.cuda.synchronize()

    def step_graph(self):
        """One step by graph replay (capture() first)."""
        if self.scheme == "dsph":
            self.graph[0].replay()
        else:
            g1, g2, g3 = self.graph
            g1.replay()
            self._dfsph_loop(graphs=g2)
            g3.replay()
        self.steps_done += 1

    # ------------------------------------------------------------ queries
    def status(self) -> np.ndarray:
        st = self.grid.status.cpu().numpy()
        if self.bgrid is not None:   # its table and the fluid -> boundary list
            st = st | self.bgrid.status.cpu().numpy()
        if self.scheme == "dfsph":
            st = st | self.df["status"].cpu().numpy()
        return st

    def check(self):
        st = self.status()
        if (st & L.ST_OVERFLOW).any():
            raise RuntimeError("pair-list capacity exceeded; rebuild the batch with a larger cap")
        # cell-walk steps build no fluid list
        nnz = (None if self.cellwalk or self.stencil_lists
               else int(self.ff.rowptr[self.NP].item()))
        return st, nnz

    def scene_state(self, k: int, extra: bool = False) -> dict:
        """State of scene k in its original particle order.  With `extra`,
        also the keys the reference leaves in simulationState after
        integrate(): `acc` (fluidAcceleration: the RK4 sum k1 + 2 k2 + 2 k3
        + k4, simulationBase.py:701, or the DFSPH step's total
        acceleration) and `p` (fluidPressure: the density solve's final
        pressure for DFSPH, dfsph.py:307; for delta-SPH the EOS of the last
        right-hand side's density, pressure.py:99-111)."""
        a, b = int(self.p_begin[k]), int(self.p_begin[k] + self.n[k])
        inv = np.empty(b - a, dtype=np.int64)
        inv[self.uid[a:b].cpu().numpy() - a] = np.arange(b - a)
        out = dict(x=self.x[a:b].cpu().numpy()[inv], v=self.v[a:b].cpu().numpy()[inv],
                   rho=self.rho[a:b].cpu().numpy()[inv])
        if extra and self.steps_done:
            out["acc"] = self.acc[a:b].cpu().numpy()[inv]
            if self.scheme == "dfsph":
                out["p"] = self.df["p"][a:b].cpu().numpy()[inv]
            else:
                # stage 4 read S[0] (the RK4 stage order of _dsph_step)
                out["p"] = eos_host(self.scenes[k].phys,
                                    self.S[0][2][a:b].cpu().numpy()[inv]).astype(np.float32)
        return out

    def dfsph_iterations(self) -> np.ndarray:
        return self.df["iters"].cpu().numpy()

    def divergence_iterations(self) -> np.ndarray:
        """Iterations of the last divergence solve per scene (zeros when the
        batch runs none)."""
        if not self.divergence:
            return np.zeros(self.nscenes, dtype=np.int32)
        return self.div_iters.cpu().numpy()


class MixedBatch:
    """delta-SPH and DFSPH scenes advanced together: each scheme (and each
    delta-plus setting of delta-SPH scenes) is one SceneBatch; the parts run
    on their own streams so their kernels overlap on the device (default
    stream priority for all)."""

    def __init__(self, scenes: list[Scene], device="cuda", carry: list | None = None,
                 scheme_kw: dict | None = None, **kw):
        """kw goes to every part; scheme_kw[scheme] to that scheme's parts
        (e.g. {"dfsph": {"cell_div": 2}})."""
        self.scenes = scenes
        self.index = {}   # scene number -> (sub-batch, position)
        self.parts = []
        carry = carry if carry is not None else [None] * len(scenes)
        def kind(s):
            return (s.scheme, bool(s.phys.get("dplus", False)), bool(s.phys.get("shifting", False)),
                    bool(s.phys.get("divergence", False)))

        for key in sorted({kind(s) for s in scenes}, key=lambda k: (k[0] != "dsph", k)):
            ids = [k for k, s in enumerate(scenes) if kind(s) == key]
            if ids:
                b = SceneBatch([scenes[k] for k in ids], device,
                               carry=[carry[k] for k in ids],
                               **{**kw, **(scheme_kw or {}).get(key[0], {})})
                for pos, k in enumerate(ids):
                    self.index[k] = (b, pos)
                self.parts.append(b)
        self.streams = [torch.cuda.Stream() for _ in self.parts]
        self.n_particles = int(sum(s.n for s in scenes))
        self.captured = False

    def capture(self, warm: bool = True) -> dict:
        """Records each part's graphs.  With `warm`, one eager step of every
        scene comes first (both parts on their own streams, as in step());
        without it no scene advances.  Returns seconds per phase."""
        t0 = time.perf_counter()
        if warm:
            self.step()
            torch.cuda.synchronize()
        t1 = time.perf_counter()
        for b in self.parts:
            b.capture(warm=False)
        self.captured = True
        return dict(eager=t1 - t0, record=time.perf_counter() - t1)

    def step(self, serial: bool = False):
        if serial:   # both parts on the current stream, one after the other
            for b in self.parts:
                if self.captured:
                    b.step_graph()
                else:
                    b.step()
            return
        cur = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(cur)
        # the DFSPH part polls the host between iteration chunks, so the
        # delta-SPH graph is queued first and runs underneath those polls
        for b, s in sorted(zip(self.parts, self.streams), key=lambda t: t[0].scheme != "dsph"):
            with torch.cuda.stream(s):
                if self.captured:
                    b.step_graph()
                else:
                    b.step(s)
        for s in self.streams:
            cur.wait_stream(s)

    def scene_state(self, k, extra: bool = False):
        b, pos = self.index[k]
        return b.scene_state(pos, extra)

    def export(self, k):
        b, pos = self.index[k]
        return b.export(pos)

    def status(self):
        return np.concatenate([b.status() for b in self.parts])
