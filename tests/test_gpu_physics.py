"""GPU kernel-sum gathers, boundary models, DFSPH solve and full steps
against the FP64 oracle.

Tolerances (fp32 device arithmetic vs FP64):
  densities                      1e-5 relative
  dpdt / accelerations           max |gpu - oracle| <= 2e-4 * max |oracle|
  mDBC boundary density          1e-4 relative
  DFSPH alpha / source           1e-4 of max
  trajectories after N steps     positions within 1e-5 * tank size
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pytorchsph_amd import _lib as L
from pytorchsph_amd import scenes as S
from pytorchsph_amd.engine import SceneBatch
from tests.gpu_helpers import assert_same_pairs, csr_numpy, in_scene_order, rel_err

pytestmark = pytest.mark.gpu


def perturbed_dsph_scene(dx=0.01, seed=0, diffusion="MOG"):
    sc = S.dam_break(dx=dx, scheme="dsph", seed=seed)
    rng = np.random.default_rng(seed + 100)
    sc.v = rng.normal(0, 0.3, sc.v.shape).astype(np.float32)
    sc.phys = dict(sc.phys, diffusion=diffusion)
    return sc


def run_rhs(sc, rho):
    """The engine's right-hand side (the cell walk for MOG scenes), with every
    pair list of the state built as well."""
    b = SceneBatch([sc])
    b.rho.copy_(torch.from_numpy(rho.astype(np.float32)))
    b.search(b.x)
    b.dsph_rhs(b.x, b.v, b.rho)
    torch.cuda.synchronize()
    return b


@pytest.mark.parametrize("diffusion", ["MOG", "deltaSPH"])
def test_dsph_rhs_with_mdbc(cuda, diffusion):
    sc = perturbed_dsph_scene(diffusion=diffusion)
    rng = np.random.default_rng(1)
    rho = (1.0 + rng.uniform(-2e-3, 4e-3, sc.n)).astype(np.float32)
    b = run_rhs(sc, rho)
    so = O.SceneOracle(sc)
    ff, bf, gf, fb = so.pairs(sc.x)
    # pair lists bit-exact
    assert_same_pairs(*csr_numpy(b.ff.rowptr, b.ff.col, sc.n), ff)
    assert_same_pairs(*csr_numpy(b.gf.rowptr, b.gf.col, sc.nb), gf)
    assert_same_pairs(*csr_numpy(b.fb.rowptr, b.fb.col, sc.n), fb)
    f = dict(x=sc.x, v=sc.v, rho=rho, area=so.area, rest=so.rest)
    rho_b, p_b, cnt = O.mdbc(f, sc.phys, dict(x=sc.bx, gx=sc.bgx, vol=sc.bvol), gf)
    assert np.array_equal(b.bcount[: sc.nb].cpu().numpy(), cnt)
    assert rel_err(b.brho[: sc.nb].cpu().numpy(), rho_b) < 1e-4
    bd = dict(x=sc.bx, rho=b.brho[: sc.nb].cpu().numpy(), vol=sc.bvol,
              p=b.bp[: sc.nb].cpu().numpy())
    dpdt, acc, _ = O.dsph_rhs(f, sc.phys, ff, bd, fb, O.dsph_terms(sc.phys))
    assert rel_err(b.dpdt.cpu().numpy(), dpdt) < 2e-4
    assert rel_err(b.acc.cpu().numpy(), acc) < 2e-4


@pytest.mark.parametrize("term", ["CONT", "MOG", "PRESS", "VISC", "LAM", "CONT_B", "PRESS_B",
                                  "VISC_B", "GRAV"])
def test_each_term_separately(cuda, term):
    """Every reference module's term on its own (momentum.py, pressure.py,
    diffusion.py, laminar.py, densityDiffusion.py and the solid boundary
    terms), through the accumulate path the module API uses."""
    sc = perturbed_dsph_scene(dx=0.02)
    rng = np.random.default_rng(2)
    rho = (1.0 + rng.uniform(-2e-3, 4e-3, sc.n)).astype(np.float32)
    b = run_rhs(sc, rho)
    bit = getattr(L, "T_" + term)
    F = b.fluid(b.x, b.v, b.rho)
    import ctypes as C
    b.dpdt.fill_(1.0)
    b.acc.fill_(1.0)
    ff, fb, B = b.ff.csr().c(), b.fb.csr().c(), b.boundary()
    L.call("sph_dsph_gather", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B), C.byref(fb),
           None, bit | L.T_ACCUM, None, L.ptr(b.dpdt), L.ptr(b.acc), L.stream_handle())
    torch.cuda.synchronize()
    so = O.SceneOracle(sc)
    off, bfo, gf, fbo = so.pairs(sc.x)
    f = dict(x=sc.x, v=sc.v, rho=rho, area=so.area, rest=so.rest)
    bd = dict(x=sc.bx, rho=b.brho[: sc.nb].cpu().numpy(), vol=sc.bvol,
              p=b.bp[: sc.nb].cpu().numpy())
    dpdt, acc, _ = O.dsph_rhs(f, sc.phys, off, bd, fbo, bit)
    if np.abs(dpdt).max() > 0:
        assert rel_err(b.dpdt.cpu().numpy() - 1.0, dpdt) < 2e-4
    if np.abs(acc).max() > 0:
        assert rel_err(b.acc.cpu().numpy() - 1.0, acc) < 2e-4


def test_fluid_and_akinci_density(cuda):
    sc = S.dam_break(dx=0.01, block=(1.0, 1.0), scheme="dfsph", seed=3)
    b = SceneBatch([sc])
    b._dfsph_pre()
    torch.cuda.synchronize()
    so = O.SceneOracle(sc)
    ff, bf, _, fb = so.pairs(sc.x)
    # the batch holds its fluid in cell order: columns and fluid arrays go
    # back to the scene's order through the particle ids
    rp, col = csr_numpy(b.bf.rowptr, b.bf.col, sc.nb)
    assert_same_pairs(rp, b.uid.cpu().numpy()[col], bf)
    f = dict(x=sc.x, v=sc.v, rho=np.ones(sc.n), area=so.area, rest=so.rest)
    bd = dict(x=sc.bx, vol=sc.bvol, rest=so.brest)
    rho, vol = O.fluid_density(f, sc.phys, ff, bd, fb)
    rho_g, vol_g = in_scene_order(b, b.rho), in_scene_order(b, b.vol)
    assert rel_err(rho_g, rho) < 1e-5
    assert np.max(np.abs(rho_g / rho - 1)) < 1e-5
    assert np.max(np.abs(vol_g / vol - 1)) < 1e-5
    brho = O.akinci_bdensity(f, sc.phys, bd, sc.bterm, bf)
    assert np.max(np.abs(b.brho[: sc.nb].cpu().numpy() / brho - 1)) < 1e-5


def test_dfsph_alpha_source_and_fixed_iterations(cuda):
    sc = S.dam_break(dx=0.01, block=(1.0, 0.6), scheme="dfsph", seed=4)
    # a step of 0.04 h, large enough that the solve runs all five iterations
    sc.phys["dt"] = 0.04 * float(sc.support)
    rng = np.random.default_rng(4)
    sc.v = rng.normal(0, 0.5, sc.v.shape).astype(np.float32)
    b = SceneBatch([sc])
    b._dfsph_pre()
    b._dfsph_iters(5)
    torch.cuda.synchronize()
    so = O.SceneOracle(sc)
    ff, bf, _, fb = so.pairs(sc.x)
    rho = in_scene_order(b, b.rho).astype(np.float64)
    brho = b.brho[: sc.nb].cpu().numpy().astype(np.float64)
    f = dict(x=sc.x, v=sc.v, rho=rho, area=so.area, rest=so.rest)
    bd = dict(x=sc.bx, vol=sc.bvol, rest=so.brest, rho=brho)
    acc = np.tile([0.0, -S.G], (sc.n, 1))
    out = O.dfsph(f, sc.phys, ff, bd, fb, bf, acc, fixed_iters=5)
    d = b.df
    assert rel_err(in_scene_order(b, d["alpha"]), out["alpha"]) < 1e-4
    assert rel_err(in_scene_order(b, d["src"]), out["src"]) < 1e-4
    assert rel_err(d["b_alpha"][: sc.nb].cpu().numpy(), out["b_alpha"]) < 1e-4
    assert rel_err(d["b_src"][: sc.nb].cpu().numpy(), out["b_src"]) < 1e-4
    assert int(d["iters"][0].item()) == 5
    hist = d["err_hist"][0, :5].cpu().numpy()
    assert np.allclose(hist, out["err_hist"][:5], rtol=1e-3, atol=1e-7)
    assert rel_err(in_scene_order(b, d["p2"]), out["p2"]) < 1e-3


def test_dfsph_converged_step(cuda):
    sc = S.dam_break(dx=0.01, block=(1.0, 0.6), scheme="dfsph", seed=5)
    b = SceneBatch([sc])
    b.step()
    torch.cuda.synchronize()
    so = O.SceneOracle(sc)
    x1, v1, rho, out = so.dfsph_step(sc.x, sc.v)
    it_gpu = int(b.df["iters"][0].item())
    assert abs(it_gpu - out["iters"]) <= 1
    st = b.scene_state(0)
    assert rel_err(st["v"], v1) < 1e-3
    assert np.max(np.abs(st["x"] - x1)) < 1e-6


@pytest.mark.parametrize("reorder", [True, False])
def test_dsph_trajectory_ten_steps(cuda, reorder):
    """Ten RK4 steps; with reorder the state is permuted into cell order at
    every step and read back through the particle ids."""
    sc = S.dam_break(dx=0.02, block=(1.0, 0.5), scheme="dsph", seed=6)
    b = SceneBatch([sc], reorder=reorder)
    so = O.SceneOracle(sc)
    x, v, r = sc.x.astype(np.float64), sc.v.astype(np.float64), np.ones(sc.n)
    for _ in range(10):
        b.step()
        x, v, r = so.dsph_step(x, v, r)
    torch.cuda.synchronize()
    st, _ = b.check()
    assert st[0] == 0
    got = b.scene_state(0)
    assert np.max(np.abs(got["x"] - x)) < 1e-5 * 4.0
    assert np.max(np.abs(got["rho"] - r)) < 1e-5
    assert rel_err(got["v"], v) < 1e-3


def test_dfsph_trajectory_three_steps(cuda):
    sc = S.dam_break(dx=0.02, block=(0.6, 0.6), scheme="dfsph", seed=7)
    b = SceneBatch([sc])
    so = O.SceneOracle(sc)
    x, v = sc.x.astype(np.float64), sc.v.astype(np.float64)
    for _ in range(3):
        b.step()
        x, v, _, _ = so.dfsph_step(x, v)
    torch.cuda.synchronize()
    got = b.scene_state(0)
    assert np.max(np.abs(got["x"] - x)) < 1e-5 * 4.0
    assert rel_err(got["v"], v) < 2e-3


def test_packed_gather_equals_structure_of_arrays(cuda):
    """sph_dsph_gather_packed and sph_dsph_gather compute the same terms."""
    import ctypes as C
    sc = perturbed_dsph_scene(dx=0.01)
    rng = np.random.default_rng(3)
    rho = (1.0 + rng.uniform(-2e-3, 4e-3, sc.n)).astype(np.float32)
    b = run_rhs(sc, rho)                       # packed path
    dp_packed, acc_packed = b.dpdt.clone(), b.acc.clone()
    F, B = b.fluid(b.x, b.v, b.rho), b.boundary()
    ff, fb = b.ff.csr().c(), b.fb.csr().c()
    L.call("sph_dsph_gather", C.byref(F), L.ptr(b.phys), C.byref(ff), C.byref(B), C.byref(fb),
           None, b.terms, None, L.ptr(b.dpdt), L.ptr(b.acc), L.stream_handle())
    torch.cuda.synchronize()
    assert rel_err(dp_packed.cpu().numpy(), b.dpdt.cpu().numpy()) < 1e-6
    assert rel_err(acc_packed.cpu().numpy(), b.acc.cpu().numpy()) < 1e-6


def test_pinv2x2_device_matches_oracle(cuda):
    rng = np.random.default_rng(5)
    M = rng.normal(size=(4096, 2, 2)).astype(np.float32)
    M[:, 0, 0] += 2
    M[:, 1, 1] += 2
    Mt = torch.from_numpy(M.reshape(-1, 4)).cuda()
    Lt = torch.empty_like(Mt)
    et = torch.empty((4096, 2), dtype=torch.float32, device="cuda")
    L.call("sph_pinv2x2", L.ptr(Mt), 4096, L.ptr(Lt), L.ptr(et), L.stream_handle())
    torch.cuda.synchronize()
    Lo, eo = O.pinv2x2(M.astype(np.float64))
    Lg = Lt.cpu().numpy().reshape(-1, 2, 2).astype(np.float64)
    # The closed form goes through the SVD angles, which are ill-conditioned
    # in fp32 when the singular values are close (eps * s1 / (s1 - s2)), and
    # gets s2 from sqrt((S1 - S2) / 2), which cancels when s2 << s1
    # (eps * (s1 / s2)^2).  The reference evaluates the same formula in fp32.
    sv = np.linalg.svd(M.astype(np.float64), compute_uv=False)
    cond = np.maximum(sv[:, 0] / np.maximum(sv[:, 0] - sv[:, 1], 1e-30),
                      (sv[:, 0] / sv[:, 1]) ** 2)
    err = np.linalg.norm(Lg - Lo, axis=(1, 2)) / np.linalg.norm(Lo, axis=(1, 2))
    bound = 50 * np.finfo(np.float32).eps * np.maximum(cond, 1.0)
    assert np.all(err <= bound), (err.max(), np.argmax(err / bound))
    assert np.median(err) < 1e-6
    # singular values: s2 = sqrt((S1 - S2) / 2) carries eps * s1^2 / s2
    de = np.abs(np.sort(et.cpu().numpy(), 1) - np.sort(eo, 1))
    tol = 50 * np.finfo(np.float32).eps * sv[:, 0] ** 2 / sv[:, 1]
    assert np.all(de <= tol[:, None])


def test_pinv2x2_device_special_cases_are_textbook(cuda):
    """The zero, rank-1 and diag(a, < 1e-5) matrices (util.py:386-433 drops
    singular values at or below 1e-5 and zeroes a singular vector with
    sign(0) = 0 only where its value is dropped) give the textbook
    pseudo-inverse on the device too, as on the oracle
    (tests/test_oracle_pins.py)."""
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    cases = [np.zeros((2, 2)), np.array([[1.0, 2.0], [2.0, 4.0]]),
             np.array([[1.0, 2.0], [0.0, 0.0]]), np.array([[0.0, 0.0], [3.0, -1.0]]),
             np.diag([2.0, 5e-6]), np.diag([3e-6, 2.0]), R @ np.diag([2.0, 4e-6]) @ R.T]
    M = np.stack(cases).astype(np.float32)
    n = len(cases)
    Mt = torch.from_numpy(M.reshape(-1, 4)).cuda()
    Lt = torch.empty_like(Mt)
    et = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    L.call("sph_pinv2x2", L.ptr(Mt), n, L.ptr(Lt), L.ptr(et), L.stream_handle())
    torch.cuda.synchronize()
    Lg = Lt.cpu().numpy().reshape(-1, 2, 2).astype(np.float64)
    eg = et.cpu().numpy()
    for k, A in enumerate(M.astype(np.float64)):
        U, sv, Vt = np.linalg.svd(A)
        inv = np.where(sv > 1e-5, 1.0 / np.where(sv > 1e-5, sv, 1.0), 0.0)
        want = Vt.T @ np.diag(inv) @ U.T
        assert np.allclose(Lg[k], want, atol=2e-6 * max(1.0, np.abs(want).max())), k
        assert abs(abs(eg[k, 0]) - sv[0]) <= 1e-6 * max(sv[0], 1.0), k
