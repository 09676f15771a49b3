"""Lanes-per-row knobs: the pair lists are bit-identical for every knob value
(including the order inside each row), and the gathers and the DFSPH solve
stay within the oracle tolerances for every value."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pytorchsph_amd import _lib as L
from pytorchsph_amd import scenes as S
from pytorchsph_amd.engine import SceneBatch
from tests.gpu_helpers import (assert_same_pairs, csr_numpy, gpu_search, rel_err, sorted_rows,
                                tuned)
from tests.test_gpu_physics import perturbed_dsph_scene
from tests.test_gpu_search import oracle_self

pytestmark = pytest.mark.gpu

LANES = [1, 2, 4, 8, 16, 32]


def test_knob_validation(cuda):
    old = L.get_tuning()
    for bad in (0, 3, 64, -1):
        with pytest.raises(RuntimeError):
            L.set_tuning(search=bad)
    assert L.get_tuning() == old


def test_search_rows_identical_across_lanes(cuda):
    x = S.point_cloud(50_000, seed=11)
    hmax = S.support_for_mean_neighbors(x, 20.0, sample=5000)
    h = np.full(len(x), hmax, np.float32)
    ref = None
    for kl in LANES:
        with tuned(search=kl, fill=kl):
            rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
        assert st == 0
        if ref is None:
            ref = (rp, c)
            assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))
        else:
            assert np.array_equal(rp, ref[0]) and np.array_equal(c, ref[1]), f"lanes {kl}"


@pytest.mark.parametrize("kl", [1, 8, 32])
def test_dsph_rhs_lanes(cuda, kl):
    sc = perturbed_dsph_scene(dx=0.01, seed=2)
    rng = np.random.default_rng(5)
    rho = (1.0 + rng.uniform(-2e-3, 4e-3, sc.n)).astype(np.float32)
    with tuned(gather=kl, search=kl, walk=kl):
        b = SceneBatch([sc])
        b.rho.copy_(torch.from_numpy(rho))
        b.dsph_rhs(b.x, b.v, b.rho)
        torch.cuda.synchronize()
    so = O.SceneOracle(sc)
    ff, _, _, fb = so.pairs(sc.x)
    f = dict(x=sc.x, v=sc.v, rho=rho, area=so.area, rest=so.rest)
    bd = dict(x=sc.bx, rho=b.brho[: sc.nb].cpu().numpy(), vol=sc.bvol,
              p=b.bp[: sc.nb].cpu().numpy())
    dpdt, acc, _ = O.dsph_rhs(f, sc.phys, ff, bd, fb, O.dsph_terms(sc.phys))
    assert rel_err(b.dpdt.cpu().numpy(), dpdt) < 2e-4
    assert rel_err(b.acc.cpu().numpy(), acc) < 2e-4


@pytest.mark.parametrize("kl", [1, 4, 16, 32])
def test_dfsph_lanes(cuda, kl):
    sc = S.dam_break(dx=0.01, block=(1.0, 0.6), scheme="dfsph", seed=4)
    # a step of 0.04 h, large enough that the solve runs all five iterations
    sc.phys["dt"] = 0.04 * float(sc.support)
    rng = np.random.default_rng(4)
    sc.v = rng.normal(0, 0.5, sc.v.shape).astype(np.float32)
    # the CSR gathers, which take the lanes-per-row knobs (the stencil-list
    # gathers run one lane per row)
    with tuned(solve=kl, gather=kl):
        b = SceneBatch([sc], stencil_lists=False)
        b._dfsph_pre()
        b._dfsph_iters(5)
        torch.cuda.synchronize()
    so = O.SceneOracle(sc)
    ff, bf, _, fb = so.pairs(sc.x)
    rho = b.rho.cpu().numpy().astype(np.float64)
    brho = b.brho[: sc.nb].cpu().numpy().astype(np.float64)
    f = dict(x=sc.x, v=sc.v, rho=rho, area=so.area, rest=so.rest)
    bd = dict(x=sc.bx, vol=sc.bvol, rest=so.brest, rho=brho)
    acc = np.tile([0.0, -S.G], (sc.n, 1))
    out = O.dfsph(f, sc.phys, ff, bd, fb, bf, acc, fixed_iters=5)
    d = b.df
    rho_o, _ = O.fluid_density(f, sc.phys, ff, bd, fb)
    assert np.max(np.abs(b.rho.cpu().numpy() / rho_o - 1)) < 1e-5
    assert rel_err(d["alpha"].cpu().numpy(), out["alpha"]) < 1e-4
    assert rel_err(d["src"].cpu().numpy(), out["src"]) < 1e-4
    assert rel_err(d["b_alpha"][: sc.nb].cpu().numpy(), out["b_alpha"]) < 1e-4
    hist = d["err_hist"][0, :5].cpu().numpy()
    assert np.allclose(hist, out["err_hist"][:5], rtol=1e-3, atol=1e-7)
    assert rel_err(d["p2"].cpu().numpy(), out["p2"]) < 1e-3


def test_knob_validation_block(cuda):
    old = L.get_tuning()
    for bad in (32, 96, 512, 0):
        with pytest.raises(RuntimeError):
            L.set_tuning(block=bad)
    assert L.get_tuning() == old


def test_workgroup_size_changes_nothing(cuda):
    """The workgroup size of the list gathers and the per-query search only
    regroups rows into workgroups: every pair list and every gathered value
    is the same bit for bit at 64, 128 and 256 threads."""
    import ctypes as C

    sc = S.dam_break(dx=0.01, scheme="dsph", seed=3)
    ref = None
    for bs in (256, 128, 64):
        with tuned(block=bs):
            b = SceneBatch([sc], cellwalk=False)
            b.search(b.x)
            ff, fbp = b._dsph_prepare(b.x, b.v, b.rho, b.frec)
            L.call("sph_dsph_gather_packed", L.ptr(b.frec), L.ptr(b.brec), L.ptr(b.pscene), b.NP,
                   L.ptr(b.phys), C.byref(ff), fbp, None, b.terms, None, L.ptr(b.dpdt),
                   L.ptr(b.acc), L.stream_handle())
            F, B = b.fluid(b.x, b.v, b.rho), b.boundary()
            fb = b.fb.csr().c()
            rho = torch.empty_like(b.rho)
            L.call("sph_fluid_density", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B),
                   C.byref(fb), None, L.ptr(rho), None, L.stream_handle())
            torch.cuda.synchronize()
            got = [t.cpu().numpy().copy() for t in (b.ff.rowptr, b.ff.col, b.gf.rowptr, b.gf.col,
                                                     b.fb.rowptr, b.fb.col, b.dpdt, b.acc, rho)]
        if ref is None:
            ref = got
        else:
            for k, (a, r) in enumerate(zip(got, ref)):
                assert np.array_equal(a.view(np.int32), r.view(np.int32)), (bs, k)


@pytest.mark.parametrize("scheme", ["dfsph", "dsph"])
def test_half_cells_same_pairs_and_trajectory(cuda, scheme):
    """Cells of half the support with a radius-2 stencil (the autotuner's
    cell-size knob): the same pair sets as the default grid (rows sorted),
    and five steps within the oracle tolerances (the order of the sums
    inside a row follows the stencil)."""
    sc = S.dam_break(dx=0.02, block=(0.8, 0.6), scheme=scheme, seed=8)
    b1, b2 = SceneBatch([sc]), SceneBatch([sc], cell_div=2)
    assert not b2.cellwalk
    for b in (b1, b2):
        b.search(b.x)
    torch.cuda.synchronize()
    # delta-SPH batches build no boundary -> fluid list (their fluid ->
    # boundary list is searched directly), DFSPH batches no ghost list
    for name in ("ff", "fb") + (("gf",) if scheme == "dsph" else ("bf",)):
        n = b1.NP if name in ("ff", "fb") else b1.NB
        p1, p2 = getattr(b1, name), getattr(b2, name)
        rp1, c1 = csr_numpy(p1.rowptr, p1.col, n)
        rp2, c2 = csr_numpy(p2.rowptr, p2.col, n)
        assert np.array_equal(rp1, rp2), name
        assert np.array_equal(sorted_rows(rp1, c1), sorted_rows(rp2, c2)), name
    so = O.SceneOracle(sc)
    x, v, r = sc.x.astype(np.float64), sc.v.astype(np.float64), np.ones(sc.n)
    for _ in range(5):
        b2.step()
        if scheme == "dsph":
            x, v, r = so.dsph_step(x, v, r)
        else:
            x, v, r, _ = so.dfsph_step(x, v)
    torch.cuda.synchronize()
    got = b2.scene_state(0)
    assert np.max(np.abs(got["x"] - x)) < 1e-5 * 4.0
    assert rel_err(got["v"], v) < 2e-3
