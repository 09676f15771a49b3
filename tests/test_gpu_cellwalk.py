"""The cell-walk delta-SPH gather (sph_dsph_cellwalk), which finds its fluid
pairs in the kernel, against the packed gather over the pair list of
sph_neighbor_fill: dp/dt and the accelerations are bit for bit equal, for
every lanes-per-row value, on wall-bounded dam breaks, colliding drops and
batches of several scenes; on crowded cells whose stencil does not fit in
LDS (the global-memory form), rows with hundreds of neighbours (hit lists
evaluated in pieces) and points outside the cell table; for any partition
of the sorted slots into workgroup runs; fused with the RK4 stage; and over
whole steps of the engine.  The list gather itself is checked against the
FP64 oracle in the physics tests, and the engine's default step (the cell
walk) against the oracle in the workload tests."""
import ctypes as C

import numpy as np
import pytest
import torch

from pytorchsph_amd import _lib as L
from pytorchsph_amd import scenes as S
from pytorchsph_amd.engine import SceneBatch, walk_segments

pytestmark = pytest.mark.gpu


def bits(t):
    return t.detach().cpu().numpy().view(np.int32)


def both_rhs(b, segs=None):
    """(dpdt, acc) of the list gather and of the cell walk on b's state:
    [list gather, walk with the boundary pairs from the list, walk with the
    boundary pairs found in the static boundary grid]."""
    sh = L.stream_handle()
    b.search(b.x)                                     # with every list
    ff, fbp = b._dsph_prepare(b.x, b.v, b.rho, b.frec)
    out = []
    for mode in ("list", "walk+fb", "walk+grid"):
        dpdt = torch.full_like(b.dpdt, float("nan"))
        acc = torch.full_like(b.acc, float("nan"))
        if mode != "list":
            sg = b.segs if segs is None else segs
            g = b.grid._grid()
            bdy = (fbp, None) if mode == "walk+fb" else (None, b._bgrid_arg())
            L.call("sph_dsph_cellwalk", C.byref(g), L.ptr(sg), int(sg.shape[0]), L.ptr(b.frec),
                   L.ptr(b.brec), L.ptr(b.phys), *bdy, b.terms, L.ptr(dpdt), L.ptr(acc), sh)
        else:
            L.call("sph_dsph_gather_packed", L.ptr(b.frec), L.ptr(b.brec), L.ptr(b.pscene),
                   b.NP, L.ptr(b.phys), C.byref(ff), fbp, None, b.terms, b._order(),
                   L.ptr(dpdt), L.ptr(acc), sh)
        out.append((dpdt, acc))
    torch.cuda.synchronize()
    return out


def assert_bitwise(out, what):
    (d0, a0) = out[0]
    assert np.isfinite(d0.cpu().numpy()).all() and np.isfinite(a0.cpu().numpy()).all(), what
    for k, (d1, a1) in enumerate(out[1:], 1):
        assert np.array_equal(bits(d0), bits(d1)), (what, k, "dpdt")
        assert np.array_equal(bits(a0), bits(a1)), (what, k, "acc")


def scene_sets():
    return {
        "dam_break": [S.dam_break(dx=0.02, scheme="dsph", seed=1)],
        "drops": [S.colliding_drops(dx=0.015, scheme="dsph", seed=2)],
        "mix": [S.dam_break(dx=0.02, scheme="dsph", seed=1),
                S.colliding_drops(dx=0.015, scheme="dsph", seed=2),
                S.dam_break(dx=0.03, block=(0.6, 0.9), scheme="dsph", seed=3)],
    }


@pytest.fixture
def lanes():
    old = L.get_tuning()
    yield
    L.set_tuning(**old)


@pytest.mark.parametrize("kl", [1, 2, 4, 8])
@pytest.mark.parametrize("kind", ["dam_break", "drops", "mix"])
def test_cellwalk_equals_list_gather(cuda, lanes, kind, kl):
    L.set_tuning(gather=kl, walk=kl)
    b = SceneBatch(scene_sets()[kind])
    assert b.cellwalk
    for _ in range(2):            # from rest and after two steps of flow
        assert_bitwise(both_rhs(b), (kind, kl))
        b.step()
        b.step()


def compress_patch(sc, frac=0.75, n=400):
    """Squeezes the n points nearest the block's centre toward it by frac
    (density x 1 / frac^2): their rows hold more than 64 candidates while
    the rows around them do not."""
    c = sc.x.mean(0)
    near = np.argsort(((sc.x - c) ** 2).sum(1))[:n]
    sc.x[near] = (c + (sc.x[near] - c) * frac).astype(np.float32)
    return sc


@pytest.mark.parametrize("kl", [1, 4])
def test_cellwalk_dense_patch(cuda, lanes, kl):
    """Rows with more candidates than the one-lane walk's mask holds, in
    chunks whose stencil still fits in LDS, beside rows that fit."""
    L.set_tuning(gather=kl, walk=kl)
    b = SceneBatch([compress_patch(S.dam_break(dx=0.01, scheme="dsph", seed=4))])
    assert_bitwise(both_rhs(b), ("dense", kl))
    rows = np.diff(b.ff.rowptr[: b.NP + 1].cpu().numpy())
    assert rows.max() > 32 and np.median(rows) < 25
    assert not b.status()[0] & L.ST_OVERFLOW


def crowded_scene():
    """C1's dam break with 1,000 points squeezed into half a cell (each of
    them has ~1,000 neighbours) and three points outside the cell table."""
    sc = S.dam_break(dx=0.01, scheme="dsph", seed=4)
    h = float(sc.support)
    rng = np.random.default_rng(7)
    inner = np.where((sc.x[:, 0] > 0.3) & (sc.x[:, 1] > 0.2))[0][:1000]
    sc.x[inner] = np.array([0.5, 0.3]) + rng.uniform(0, 0.5 * h, (len(inner), 2))
    sc.x[inner[:3]] = [[50.0, 0.3], [0.5, -40.0], [-7.0, 30.0]]
    return sc


@pytest.mark.parametrize("kl", [1, 4])
def test_cellwalk_crowded_cells_and_outside_points(cuda, lanes, kl):
    L.set_tuning(gather=kl, walk=kl)
    b = SceneBatch([crowded_scene()], pair_cap=300.0)
    out = both_rhs(b)
    st, nnz = b.status(), int(b.ff.rowptr[b.NP].item())
    assert st[0] & L.ST_OUTSIDE and not st[0] & L.ST_OVERFLOW
    rows = np.diff(b.ff.rowptr[: b.NP + 1].cpu().numpy())
    # the crowded rows have far more hits than a row's list in LDS holds,
    # and their cells far more points than a run's stencil can stage
    assert rows.max() > 900 and nnz > 900_000
    # the outside points have no fluid neighbour but themselves
    (d0, a0) = out[0]
    for d1, a1 in out[1:]:
        assert np.array_equal(bits(d0), bits(d1)) and np.array_equal(bits(a0), bits(a1))


def test_cellwalk_any_partition(cuda, lanes):
    L.set_tuning(gather=4, walk=4)
    b = SceneBatch(scene_sets()["mix"])
    ref = both_rhs(b)
    dev = b.x.device
    parts = {
        "runs of 1000": walk_segments(b.p_begin, b.n, 1000),
        "runs of 37": walk_segments(b.p_begin, b.n, 37),
        "one run over every scene": np.array([[0, b.NP]], np.int32),
        "with empty runs": np.concatenate([walk_segments(b.p_begin, b.n, 256),
                                           np.array([[5, 5], [b.NP, b.NP]], np.int32)]),
    }
    for what, sg in parts.items():
        out = both_rhs(b, torch.from_numpy(np.ascontiguousarray(sg)).to(dev))
        assert_bitwise([ref[0], *out[1:]], what)


@pytest.mark.parametrize("kl", [1, 4])
@pytest.mark.parametrize("fuse", [False, True])
def test_cellwalk_steps_equal_list_steps(cuda, lanes, fuse, kl):
    L.set_tuning(gather=kl, walk=kl)
    scenes = scene_sets()["mix"]
    walk = SceneBatch(scenes, fuse_rk4=fuse)
    lst = SceneBatch(scenes, fuse_rk4=fuse, cellwalk=False)
    assert walk.cellwalk and not lst.cellwalk
    walk.capture()
    lst.step()
    for _ in range(4):
        walk.step_graph()
        lst.step()
    torch.cuda.synchronize()
    assert (walk.status() == lst.status()).all()
    for k in range(len(scenes)):
        a, c = walk.scene_state(k), lst.scene_state(k)
        for key in ("x", "v", "rho"):
            assert np.array_equal(a[key], c[key]), (k, key, fuse)


def test_cellwalk_rejects_ddsph(cuda):
    b = SceneBatch(scene_sets()["dam_break"])
    g = b.grid._grid()
    with pytest.raises(RuntimeError, match="DDSPH"):
        L.call("sph_dsph_cellwalk", C.byref(g), L.ptr(b.segs), b.nseg, L.ptr(b.frec),
               L.ptr(b.brec), L.ptr(b.phys), None, None, L.T_DDSPH, L.ptr(b.dpdt),
               L.ptr(b.acc), L.stream_handle())


@pytest.mark.parametrize("kind", ["dam_break", "mix"])
def test_mdbc_walk_records_equal_pack(cuda, kind):
    """sph_mdbc_density_walk_records: the boundary densities, pressures and
    counts of sph_mdbc_density_walk, and the boundary records that
    sph_dsph_pack writes from them, bit for bit (walls at rest and, with a
    velocity array, moving)."""
    b = SceneBatch(scene_sets()[kind])
    sh = L.stream_handle()
    bv = torch.randn((max(b.NB, 1), 2), dtype=torch.float32, device="cuda")
    for _ in range(2):
        b.search(b.x, fluid_pairs=False, ghost_pairs=False, bdy_pairs=False)
        F = b.fluid(b.x, b.v, b.rho)
        g = b.grid._grid()
        for vel in (None, bv):
            rb, pb = torch.full_like(b.brho, float("nan")), torch.full_like(b.bp, float("nan"))
            cnt = torch.full_like(b.bcount, -1)
            L.call("sph_mdbc_density_walk", C.byref(F), L.ptr(b.phys), L.ptr(b.bx),
                   L.ptr(b.bgx), L.ptr(b.bh), L.ptr(b.bscene), b.NB, C.byref(g), L.ptr(rb),
                   L.ptr(pb), L.ptr(cnt), sh)
            Bd = L.SphBoundary(L.ptr(b.bx), L.ptr(vel) if vel is not None else None, L.ptr(rb),
                               L.ptr(b.bvol), L.ptr(pb), L.ptr(b.brest), L.ptr(b.bscene), b.NB)
            ref = torch.full_like(b.brec, float("nan"))
            L.call("sph_dsph_pack", C.byref(F), L.ptr(b.phys), C.byref(Bd), None, L.ptr(ref), sh)
            rb2, pb2 = torch.full_like(rb, float("nan")), torch.full_like(pb, float("nan"))
            cnt2 = torch.full_like(cnt, -1)
            rec = torch.full_like(b.brec, float("nan"))
            L.call("sph_mdbc_density_walk_records", C.byref(F), L.ptr(b.phys), L.ptr(b.bx),
                   L.ptr(vel) if vel is not None else None, L.ptr(b.bvol), L.ptr(b.bgx),
                   L.ptr(b.bh), L.ptr(b.bscene), b.NB, C.byref(g), L.ptr(rb2), L.ptr(pb2),
                   L.ptr(cnt2), L.ptr(rec), sh)
            torch.cuda.synchronize()
            n = b.NB
            assert np.array_equal(bits(rb[:n]), bits(rb2[:n]))
            assert np.array_equal(bits(pb[:n]), bits(pb2[:n]))
            assert np.array_equal(cnt[:n].cpu().numpy(), cnt2[:n].cpu().numpy())
            assert np.array_equal(bits(ref[:n]), bits(rec[:n]))
        for _ in range(3):
            b.step()


@pytest.mark.parametrize("kind", ["dam_break", "drops", "mix"])
def test_mdbc_walk_equals_list(cuda, kind):
    """sph_mdbc_density_walk (each ghost's fluid neighbours found in the
    kernel) against sph_mdbc_density over the ghost list of
    sph_neighbor_fill: boundary densities, pressures and neighbour counts
    bit for bit, from rest and after the fluid has moved."""
    b = SceneBatch(scene_sets()[kind])
    sh = L.stream_handle()
    for _ in range(2):
        b.search(b.x)                             # with the ghost list
        F = b.fluid(b.x, b.v, b.rho)
        gf, g = b.gf.csr().c(), b.grid._grid()
        outs = []
        for walk in (False, True):
            rb = torch.full_like(b.brho, float("nan"))
            pb = torch.full_like(b.bp, float("nan"))
            cnt = torch.full_like(b.bcount, -1)
            if walk:
                L.call("sph_mdbc_density_walk", C.byref(F), L.ptr(b.phys), L.ptr(b.bx),
                       L.ptr(b.bgx), L.ptr(b.bh), L.ptr(b.bscene), b.NB, C.byref(g), L.ptr(rb),
                       L.ptr(pb), L.ptr(cnt), sh)
            else:
                L.call("sph_mdbc_density", C.byref(F), L.ptr(b.phys), L.ptr(b.bx), L.ptr(b.bgx),
                       L.ptr(b.bscene), b.NB, C.byref(gf), L.ptr(rb), L.ptr(pb), L.ptr(cnt), sh)
            outs.append((rb[: b.NB], pb[: b.NB], cnt[: b.NB]))
        torch.cuda.synchronize()
        (r0, p0, c0), (r1, p1, c1) = outs
        # the drops touch no wall: their ghosts see no fluid
        if kind != "drops":
            assert int((c0 > 10).sum()) > 0 and int((c0 == 0).sum()) > 0
        assert np.array_equal(c0.cpu().numpy(), c1.cpu().numpy())
        assert np.array_equal(bits(r0), bits(r1)) and np.array_equal(bits(p0), bits(p1))
        for _ in range(3):
            b.step()
