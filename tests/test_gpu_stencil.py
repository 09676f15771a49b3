"""Stencil lists (sph_stencil_list_build) and the DFSPH gathers over them
(the sph_*_sl entry points).  The masks decode to the rows of
sph_neighbor_fill in its order -- on cell-ordered and randomly ordered
points, several scenes, and crowded cells whose runs are longer than a mask
(those keep their first 64 candidates and flag SPH_ST_OVERFLOW) -- and
DFSPH batches stepping over stencil lists equal batches stepping over CSR
lists with one lane per row, bit for bit: with and without walls, with the
divergence solve, eagerly and from captured graphs.  The CSR gathers
themselves are checked against the FP64 oracle in the physics and workload
tests."""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_helpers import csr_numpy, tuned
from pytorchsph_amd import _lib as L
from pytorchsph_amd import scenes as S
from pytorchsph_amd.engine import SceneBatch
from pytorchsph_amd.search import NeighborGrid

pytestmark = pytest.mark.gpu


def decode(runs, masks):
    """Rows of sorted slots from a stencil list: [n][192] slot table and
    validity, in run order then bit order."""
    n = runs.shape[0]
    bit = np.arange(64, dtype=np.uint64)
    slots, ok = [], []
    for q in range(3):
        a = runs[:, q].astype(np.int64)
        ln = (runs[:, 3].astype(np.int64) >> (8 * q)) & 255
        m = masks[:, q].view(np.uint64)
        on = ((m[:, None] >> bit[None, :]) & np.uint64(1)).astype(bool)
        on &= np.arange(64)[None, :] < np.minimum(ln, 64)[:, None]
        slots.append(a[:, None] + np.arange(64)[None, :])
        ok.append(on)
    return np.concatenate(slots, 1), np.concatenate(ok, 1)


def check_against_fill(x, h, pscene, scenes):
    """Builds the grid, the stencil list and the fill's CSR of the points
    and compares them row by row (the list's slots mapped to points)."""
    dev = torch.device("cuda")
    n = len(x)
    g = NeighborGrid(scenes, n, dev)
    xt = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    ht = torch.from_numpy(np.ascontiguousarray(h, np.float32)).to(dev)
    ps = torch.from_numpy(np.ascontiguousarray(pscene, np.int32)).to(dev)
    g.build(xt, ps)
    csr = g.search(xt, ht, ps, 1, self_sorted=True)
    runs = torch.full((n, 4), -7, dtype=torch.int32, device=dev)
    masks = torch.zeros((3, n), dtype=torch.int64, device=dev)
    gc = g._grid()
    L.call("sph_stencil_list_build", C.byref(gc), L.ptr(ht), L.ptr(runs), L.ptr(masks),
           L.stream_handle())
    torch.cuda.synchronize()
    rp, col = csr_numpy(csr.rowptr, csr.col, csr.nrows)
    sidx = g.sorted_idx[:n].cpu().numpy().astype(np.int64)
    slots, ok = decode(runs.cpu().numpy(), masks.cpu().numpy().T.copy())
    over = (np.stack([(runs[:, 3].cpu().numpy() >> (8 * q)) & 255 for q in range(3)], 1) > 64)
    for r in range(n):
        p = sidx[r]
        got = sidx[slots[r][ok[r]]]
        want = col[rp[p]:rp[p + 1]]
        if over[r].any():
            # the first 64 candidates of each long run: a subsequence
            assert set(got.tolist()) <= set(want.tolist()), r
        else:
            assert np.array_equal(got, want), (r, p)
    return g, over


def lattice(nx, ny, dx, jitter, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx)
    pts = np.stack([xs.ravel(), ys.ravel()], 1) + rng.uniform(-jitter, jitter, (nx * ny, 2)) * dx
    return pts.astype(np.float32)


@pytest.mark.parametrize("order", ["cells", "random"])
def test_stencil_list_rows_equal_fill(cuda, order):
    dx = 0.01
    pts = lattice(70, 50, dx, 0.3, 5)
    if order == "random":
        pts = pts[np.random.default_rng(1).permutation(len(pts))]
    h = np.full(len(pts), np.float32(dx * np.sqrt(20 / np.pi)), np.float32)
    sc = [dict(vmin=pts.min(0) - 0.01, vmax=pts.max(0) + 0.01, hmax=float(h[0]), p_begin=0,
               p_count=len(pts))]
    g, over = check_against_fill(pts, h, np.zeros(len(pts), np.int32), sc)
    assert not over.any() and int(g.status.cpu()[0]) & L.ST_OVERFLOW == 0


def test_stencil_list_two_scenes_and_crowded_runs(cuda):
    """Two scenes of different supports; in the second a cluster of 300
    points in one cell makes runs longer than a mask."""
    dx = 0.01
    a = lattice(40, 30, dx, 0.2, 2)
    b = lattice(30, 30, 1.5 * dx, 0.2, 3) + np.float32(5.0)
    rng = np.random.default_rng(4)
    b[:300] = np.float32(5.2) + rng.uniform(0, 0.004, (300, 2)).astype(np.float32)
    ha = np.float32(dx * np.sqrt(20 / np.pi))
    hb = np.float32(1.5 * dx * np.sqrt(20 / np.pi))
    x = np.concatenate([a, b])
    h = np.concatenate([np.full(len(a), ha), np.full(len(b), hb)]).astype(np.float32)
    ps = np.concatenate([np.zeros(len(a)), np.ones(len(b))]).astype(np.int32)
    sc = [dict(vmin=a.min(0) - 0.01, vmax=a.max(0) + 0.01, hmax=float(ha), p_begin=0,
               p_count=len(a)),
          dict(vmin=b.min(0) - 0.01, vmax=b.max(0) + 0.01, hmax=float(hb), p_begin=len(a),
               p_count=len(b))]
    g, over = check_against_fill(x, h, ps, sc)
    st = g.status.cpu().numpy()
    assert over.any() and st[1] & L.ST_OVERFLOW and not st[0] & L.ST_OVERFLOW


def dfsph_sets():
    return {
        "dam_break": [S.dam_break(dx=0.02, scheme="dfsph", seed=1)],
        "mix": [S.dam_break(dx=0.02, scheme="dfsph", seed=1),
                S.colliding_drops(dx=0.015, scheme="dfsph", seed=2),
                S.dam_break(dx=0.03, block=(0.6, 0.9), scheme="dfsph", seed=3)],
    }


def divergence_variant(scenes):
    import dataclasses
    return [dataclasses.replace(s, phys=dict(s.phys, divergence=True)) for s in scenes]


@pytest.mark.parametrize("kind", ["dam_break", "mix", "mix_divergence"])
@pytest.mark.parametrize("graphs", [False, True])
def test_stencil_list_steps_equal_csr_steps(cuda, kind, graphs):
    scenes = (divergence_variant(dfsph_sets()["mix"]) if kind == "mix_divergence"
              else dfsph_sets()[kind])
    with tuned(gather=1, solve=1):
        sl = SceneBatch(scenes, stencil_lists=True)
        cs = SceneBatch(scenes, stencil_lists=False)
        assert sl.stencil_lists and not cs.stencil_lists
        if graphs:
            sl.capture()
            cs.capture()
        for _ in range(4):
            if graphs:
                sl.step_graph()
                cs.step_graph()
            else:
                sl.step()
                cs.step()
        torch.cuda.synchronize()
        assert (sl.status() == cs.status()).all()
        for k in range(len(scenes)):
            a, c = sl.scene_state(k, extra=True), cs.scene_state(k, extra=True)
            for key in ("x", "v", "rho", "acc", "p"):
                assert np.isfinite(a[key]).all(), (k, key)
                assert np.array_equal(a[key].view(np.int32), c[key].view(np.int32)), (k, key)


def test_stencil_list_dense_patch_equals_csr(cuda):
    """A patch squeezed to 1.8x the density: its rows' three runs hold more
    than 64 candidates together (walked run after run) but at most 64 each
    (no overflow), beside rows walked as one stream."""
    from test_gpu_cellwalk import compress_patch
    sc = compress_patch(S.dam_break(dx=0.02, scheme="dfsph", seed=1), n=300)
    with tuned(gather=1, solve=1):
        sl = SceneBatch([sc], stencil_lists=True)
        cs = SceneBatch([sc], stencil_lists=False)
        sl.step()
        cs.step()
        torch.cuda.synchronize()
        runs = sl.sl_runs.cpu().numpy()
        lens = np.stack([(runs[:, 3] >> (8 * q)) & 255 for q in range(3)], 1)
        assert lens.sum(1).max() > 64 and lens.max() <= 64
        assert (np.asarray(sl.status()) == np.asarray(cs.status())).all()
        assert not sl.status()[0] & L.ST_OVERFLOW
        a, c = sl.scene_state(0, extra=True), cs.scene_state(0, extra=True)
        for key in ("x", "v", "rho", "acc", "p"):
            assert np.isfinite(a[key]).all(), key
            assert np.array_equal(a[key].view(np.int32), c[key].view(np.int32)), key
