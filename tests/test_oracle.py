"""CPU tier: pins the oracle against brute force, closed forms, numpy and
physical invariants (no GPU needed)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pytorchsph_amd import scenes as S
from pytorchsph_amd.search import grid_params


def pairset(csr):
    r, c = csr.pairs()
    return set(zip(r.tolist(), c.tolist()))


# ------------------------------------------------------------- search
def test_grid_params_match_c():
    rng = np.random.default_rng(5)
    for _ in range(200):
        vmin = rng.uniform(-3, 1, 2)
        vmax = vmin + rng.uniform(0.1, 5, 2)
        h = np.float32(rng.uniform(1e-3, 0.2))
        a = grid_params(vmin, vmax, h)
        b = O.grid_params(vmin.astype(np.float32), vmax.astype(np.float32), h)
        assert (a[0], a[1], a[2], a[3]) == (b[0], b[1], b[2], b[3])


@pytest.mark.parametrize("variable_h", [False, True])
def test_grid_search_equals_brute_force(variable_h):
    rng = np.random.default_rng(11)
    n = 2500
    x = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    hmax = np.float32(0.04)
    h = np.full(n, hmax, np.float32)
    if variable_h:
        h = rng.uniform(0.5, 1.0, n).astype(np.float32) * hmax
    qx, qy, cx, cy, _ = grid_params((0, 0), (1, 1), hmax)
    grid = O.search(x, h, x, (qx, qy), hmax, cx, 1, self_mode=True)
    brute = O.brute(x, h, x)
    assert pairset(grid) == pairset(brute)
    # foreign queries against the same grid
    q = rng.uniform(-0.05, 1.05, (700, 2)).astype(np.float32)
    hq = np.full(len(q), hmax, np.float32)
    g2 = O.search(q, hq, x, (qx, qy), hmax, cx, 1)
    b2 = O.brute(q, hq, x)
    assert pairset(g2) == pairset(b2)


def test_search_invariants():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, (1500, 2)).astype(np.float32)
    h = np.full(len(x), 0.05, np.float32)
    qx, qy, cx, _, _ = grid_params((0, 0), (1, 1), h[0])
    a = O.search(x, h, x, (qx, qy), h[0], cx, 1, self_mode=True)
    ps = pairset(a)
    assert all((i, i) in ps for i in range(len(x)))        # self pairs
    assert all((j, i) in ps for (i, j) in ps)              # symmetric for uniform h
    for r in range(a.nrows):                               # rows sorted
        seg = a.col[a.rowptr[r]:a.rowptr[r + 1]]
        assert np.all(np.diff(seg) > 0)


def test_near_tie_pairs_follow_fp32_sqrt():
    # points at distances within one ulp of h: the fp32 predicate decides
    h = np.float32(0.0625)
    base = np.array([[0.5, 0.5]], np.float32)
    others = []
    for k in range(-3, 4):
        d = np.float32(h)
        for _ in range(abs(k)):
            d = np.nextafter(d, np.float32(1 if k > 0 else 0))
        others.append([0.5 + d, 0.5])
        others.append([0.5, 0.5 - d])
    x = np.concatenate([base, np.array(others, np.float32)])
    hv = np.full(len(x), h, np.float32)
    qx, qy, cx, _, _ = grid_params((0, 0), (1, 1), h)
    got = pairset(O.search(x, hv, x, (qx, qy), h, cx, 1, self_mode=True))
    want = pairset(O.brute(x, hv, x))
    assert got == want
    for j in range(1, len(x)):
        dx = np.float32(x[0, 0] - x[j, 0])
        dy = np.float32(x[0, 1] - x[j, 1])
        inside = np.sqrt(np.float32(dx * dx + dy * dy)) < h
        assert ((0, j) in got) == bool(inside)


def sq_threshold_np(h):
    """numpy statement of the device predicate threshold (sph_common.h)."""
    h = np.float32(h)
    hp = np.nextafter(h, np.float32(0))
    m = (np.float64(hp) + np.float64(h)) * 0.5
    m2 = m * m
    f = np.float32(m2)
    if np.float64(f) > m2:
        f = np.nextafter(f, np.float32(0))
    if np.float64(f) == m2 and (int(np.array(hp).view(np.uint32)) & 1):
        f = np.nextafter(f, np.float32(0))
    return f


def test_sqrt_free_predicate_is_exact():
    """fl(sqrt(s)) < h  <=>  s <= t(h) for floats s >= 0 (the identity the
    device neighbour predicate relies on), checked against numpy's correctly
    rounded float32 sqrt around the threshold and at random."""
    rng = np.random.default_rng(0)
    hs = np.concatenate([rng.uniform(1e-4, 10, 3000).astype(np.float32),
                         np.float32(2.0) ** np.arange(-10, 4).astype(np.float32)])
    for h in hs:
        t = sq_threshold_np(h)
        cand = [t, np.nextafter(t, np.float32(0)), np.nextafter(t, np.float32(np.inf)),
                np.float32(h) * np.float32(h), np.float32(0)]
        cand += list(rng.uniform(0, 2 * float(h) ** 2, 8).astype(np.float32))
        for s in cand:
            s = np.float32(s)
            assert (np.sqrt(s) < h) == (s <= t), (h, s, t)


def test_tie_case_rounding():
    # h with an exactly representable midpoint square: construct
    # h = 1 + k ulp, hp = h - ulp, m = h - ulp/2 with m^2 exact in float.
    found = 0
    for bits in range(0x3F800000, 0x3F800000 + 4000):
        h = np.array(bits, np.uint32).view(np.float32)[()]
        hp = np.nextafter(h, np.float32(0))
        m = (np.float64(hp) + np.float64(h)) / 2
        m2 = m * m
        if np.float64(np.float32(m2)) == m2:
            found += 1
            s = np.float32(m2)
            assert (np.sqrt(s) < h) == (s <= sq_threshold_np(h))
    assert found >= 0  # ties are rare; the loop checks every one it meets


# ---------------------------------------------------------- kernels
def test_wendland_normalised_and_gradient():
    lib = O.lib()
    h = 0.37
    r = np.linspace(0, h, 20001)
    w = np.array([lib.orc_W(ri / h, h) for ri in r])
    integral = np.trapezoid(w * 2 * np.pi * r, r)
    assert abs(integral - 1.0) < 1e-6
    # grad magnitude equals -dW/dr (finite differences)
    for q in (0.1, 0.3, 0.5, 0.8):
        eps = 1e-7
        dwdr = (lib.orc_W(q + eps, h) - lib.orc_W(q - eps, h)) / (2 * eps * h)
        assert abs(lib.orc_dW(q, h) + dwdr) < 1e-5 * abs(dwdr)


def test_spiky_gradient_is_derivative_of_normalised_spiky():
    lib = O.lib()
    h = 0.21
    r = np.linspace(0, h, 20001)
    base = (1 - r / h) ** 3
    C = 1.0 / np.trapezoid(base * 2 * np.pi * r, r)      # normalisation of (1-q)^3
    for q in (0.05, 0.4, 0.9):
        want = C * 3 * (1 - q) ** 2 / h                    # -dW/dr
        assert abs(lib.orc_dWs(q, h) - want) < 1e-5 * want


# ------------------------------------------------------------- pinv
def test_pinv3x3_matches_numpy():
    rng = np.random.default_rng(2)
    for _ in range(200):
        A = rng.normal(size=(3, 3)) * rng.uniform(0.1, 10, (3, 1))
        X = O.pinv3x3(A)
        assert np.allclose(X, np.linalg.pinv(A), rtol=1e-9, atol=1e-10)
    # rank deficient
    A = np.outer([1.0, 2.0, 3.0], [4.0, -1.0, 0.5])
    assert np.allclose(O.pinv3x3(A), np.linalg.pinv(A), atol=1e-12)


def test_pinv2x2_matches_numpy_when_well_conditioned():
    rng = np.random.default_rng(4)
    M = rng.normal(size=(500, 2, 2))
    M[:, 0, 0] += 3.0
    M[:, 1, 1] += 3.0
    L, eig = O.pinv2x2(M)
    want = np.linalg.pinv(M)
    assert np.allclose(L, want, rtol=1e-9, atol=1e-12)
    sv = np.linalg.svd(M, compute_uv=False)
    assert np.allclose(np.sort(np.abs(eig), 1), np.sort(sv, 1), rtol=1e-9)


def test_pinv2x2_singular_value_cutoff():
    # singular values below 1e-5 are dropped (util.py:424-425)
    M = np.array([[[2.0, 0.0], [0.0, 5e-6]]])
    L, _ = O.pinv2x2(M)
    assert np.allclose(L[0], [[0.5, 0.0], [0.0, 0.0]], atol=1e-12)


# ------------------------------------------------------ physics invariants
def lattice(n=40, dx=0.025, jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    x = np.stack([(I.ravel() + 0.5) * dx, (J.ravel() + 0.5) * dx], 1)
    x = x + rng.uniform(-jitter * dx, jitter * dx, x.shape)
    return x.astype(np.float32)


def setup_fluid(x, dx, rho=None, v=None):
    n = len(x)
    h = S.support_for(dx)
    phys = S.default_phys("dsph", float(h))
    qx, qy, cx, _, _ = grid_params((-1, -1), (2, 2), h)
    ff = O.search(x, np.full(n, h, np.float32), x, (qx, qy), h, cx, 1, self_mode=True)
    f = dict(x=x, v=np.zeros((n, 2)) if v is None else v, rho=np.ones(n) if rho is None else rho,
             area=np.full(n, dx * dx), rest=np.full(n, 1000.0))
    interior = np.all((x > 4 * h) & (x < x.max() - 4 * h), axis=1)
    return f, phys, ff, interior


def test_summation_density_of_lattice_is_one():
    dx = 0.02
    x = lattice(50, dx)
    f, phys, ff, inner = setup_fluid(x, dx)
    rho, vol = O.fluid_density(f, phys, ff)
    assert np.all(np.abs(rho[inner] - 1.0) < 0.02)
    assert np.allclose(vol, dx * dx / rho)


def test_continuity_zero_for_translation_and_rotation():
    dx = 0.02
    x = lattice(40, dx, jitter=0.05)
    n = len(x)
    for v in (np.tile([0.3, -1.2], (n, 1)), np.stack([-x[:, 1], x[:, 0]], 1) * 2.0):
        f, phys, ff, _ = setup_fluid(x, dx, v=v.astype(np.float64))
        dpdt, _, _ = O.dsph_rhs(f, phys, ff, terms=O.T_CONT)
        assert np.max(np.abs(dpdt)) < 1e-7


def test_continuity_recovers_divergence():
    dx = 0.02
    x = lattice(50, dx)
    a = 0.7
    v = np.stack([a * x[:, 0], np.zeros(len(x))], 1)
    f, phys, ff, inner = setup_fluid(x, dx, v=v)
    dpdt, _, _ = O.dsph_rhs(f, phys, ff, terms=O.T_CONT)
    # d rho / dt = -rho div v (the lattice sum sum_j V_j x_ji (x) gradW_ij is
    # 0.97 I at this resolution, hence the 5 % band)
    assert np.allclose(dpdt[inner], -1000.0 * a, rtol=0.05)


def test_pressure_force_is_antisymmetric():
    dx = 0.02
    x = lattice(30, dx, jitter=0.2, seed=1)
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.99, 1.01, len(x))
    f, phys, ff, _ = setup_fluid(x, dx, rho=rho)
    _, acc, _ = O.dsph_rhs(f, phys, ff, terms=O.T_PRESS)
    m = dx * dx * 1000.0
    total = (m * acc).sum(0)
    assert np.all(np.abs(total) < 1e-9 * np.abs(m * acc).sum())


def test_laminar_viscosity_recovers_laplacian():
    dx = 0.02
    x = lattice(50, dx)
    v = np.stack([x[:, 1] ** 2, np.zeros(len(x))], 1)
    f, phys, ff, inner = setup_fluid(x, dx, v=v)
    _, acc, _ = O.dsph_rhs(f, phys, ff, terms=O.T_LAM)
    # a = nu lap v = (2 nu, 0); the eps = 0.1 h^2 regulariser biases it low,
    # so the magnitude is checked with eps -> 0 and the sign with the default
    assert np.all(acc[inner, 0] > 0) and np.all(acc[inner, 0] < 2 * phys["nu"])
    _, acc0, _ = O.dsph_rhs(f, dict(phys, eps=1e-12 * phys["h"] ** 2), ff, terms=O.T_LAM)
    assert np.allclose(acc0[inner, 0], 2 * phys["nu"], rtol=0.05)
    assert np.allclose(acc0[inner, 1], 0.0, atol=1e-3 * phys["nu"])


def test_renormalised_gradient_of_linear_density():
    """<grad rho> (densityDiffusion.py:45-61) of rho = rho0 (1 + c y) is
    -2 grad rho: the reference's sign and factor 2 kept literally."""
    dx = 0.02
    x = lattice(50, dx, jitter=0.1, seed=2)
    c = 0.01
    rho = 1.0 + c * x[:, 1].astype(np.float64)
    f, phys, ff, inner = setup_fluid(x, dx, rho=rho)
    _, _, grho = O.dsph_rhs(f, phys, ff, terms=O.T_DDSPH)
    assert np.allclose(grho[inner, 1], -2 * 1000.0 * c, rtol=1e-3)
    assert np.allclose(grho[inner, 0], 0.0, atol=1e-3 * 2000 * c)


def test_mdbc_density_of_fluid_at_rest():
    sc = S.dam_break(dx=0.02, block=(1.0, 0.5), scheme="dsph", seed=0)
    so = O.SceneOracle(sc)
    ff, bf, gf, fb = so.pairs(sc.x)
    f = dict(x=sc.x, v=sc.v, rho=np.ones(sc.n), area=so.area, rest=so.rest)
    rho_b, p_b, cnt = O.mdbc(f, sc.phys, dict(x=sc.bx, gx=sc.bgx, vol=sc.bvol), gf)
    assert np.all(rho_b >= 1.0)
    wet = cnt > 10
    assert wet.sum() > 50
    assert np.all(rho_b[wet] < 1.05)
    assert np.all(rho_b[cnt == 0] == 1.0)


def test_dfsph_solve_reduces_compression():
    sc = S.dam_break(dx=0.02, block=(0.5, 0.5), scheme="dfsph", seed=0)
    so = O.SceneOracle(sc)
    x1, v1, rho, out = so.dfsph_step(sc.x, sc.v)
    assert out["iters"] >= 2
    h = out["err_hist"]
    assert h[-1] <= sc.phys["tau"] or out["iters"] > sc.phys["max_iters"]
    assert np.all(out["alpha"] <= -1e-7) and np.all(out["alpha"] >= -1)
    assert np.all(out["p2"] >= 0)


def test_dsph_step_conserves_mass_and_stays_finite():
    sc = S.dam_break(dx=0.02, block=(0.5, 0.3), scheme="dsph", seed=0)
    so = O.SceneOracle(sc)
    x, v, r = sc.x.astype(np.float64), sc.v.astype(np.float64), np.ones(sc.n)
    for _ in range(3):
        x, v, r = so.dsph_step(x, v, r)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    assert np.all(np.abs(r - 1) < 0.05)
    # gravity acts downward
    assert v[:, 1].mean() < 0


def test_scene_sizes():
    sc = S.dam_break(dx=0.01, scheme="dsph")
    assert sc.n == 5000
    assert sc.nb == 406 * 206 - 400 * 200
    assert np.float32(sc.support) == S.support_for(0.01)
    pts = S.point_cloud(10000, seed=1)
    assert pts.shape == (10000, 2) and pts.min() >= 0 and pts.max() < 1
