"""Pins of the oracle's delta-plus restatement (deltaPlus.py:81-251,
pressure.py:40-52) by what the mathematics fixes, on a square lattice block
with a free surface on every side:

- lambda is the smaller singular value of M (numpy SVD);
- M -> I in the bulk (the first-moment identity of a normalised kernel), so
  lambda is near 1 there and near 1/2 on a flat free surface;
- the normal -grad(lambda)/|.| is the outward unit normal of each face;
- the surface classes follow from the one-ring spread of the face rows;
- the shift vanishes on a uniform lattice and pushes a displaced particle
  back toward its lattice site;
- the delta-plus pressure term equals the plain pressure gradient wherever
  the switch selects p_i + p_j, and vanishes for a uniform negative
  pressure in the bulk class.
"""
import numpy as np
import pytest

from oracle import oracle as O

DX = 0.01
NX, NY = 40, 30


@pytest.fixture(scope="module")
def block():
    I, J = np.meshgrid(np.arange(NX), np.arange(NY), indexing="xy")
    I, J = I.ravel(), J.ravel()
    x = np.stack([(I + 0.5) * DX, (J + 0.5) * DX], 1).astype(np.float32)
    n = len(x)
    area = DX * DX
    h = np.float32(np.sqrt(area / np.pi * 20))
    phys = dict(h=float(h), rho0=1000.0, c0=20.0, gamma=7.0, eps=0.1 * float(h) ** 2, eos=0)
    hh = np.full(n, h, np.float32)
    ff = O.brute(x, hh, x)
    f = dict(x=x.astype(np.float64), v=np.tile([1.0, 0.0], (n, 1)), rho=np.ones(n),
             area=np.full(n, area))
    M, Lm, e = O.fluid_norm(f, phys, ff)
    grad, nrm = O.dplus_normal(f, phys, ff, Lm, e)
    cmax, mask, surf = O.dplus_surface(f, phys, ff, nrm, e)
    d_edge = np.minimum.reduce([x[:, 0], x[:, 1], NX * DX - x[:, 0], NY * DX - x[:, 1]])
    return dict(I=I, J=J, x=x, h=float(h), hh=hh, phys=phys, ff=ff, f=f, M=M, L=Lm, eig=e,
                normal=nrm, mask=mask, surf=surf, d_edge=d_edge)


def test_lambda_is_smallest_singular_value(block):
    sv = np.linalg.svd(block["M"], compute_uv=False)
    assert np.allclose(block["eig"][:, 1], sv[:, 1], atol=1e-12)
    assert np.allclose(block["eig"][:, 0], sv[:, 0], atol=1e-12)
    L_np = np.linalg.pinv(block["M"])
    assert np.allclose(block["L"], L_np, atol=1e-9)


def test_lambda_bulk_and_surface(block):
    lam = block["eig"][:, 1]
    bulk = block["d_edge"] > block["h"] + DX
    assert np.all(np.abs(lam[bulk] - 1) < 0.05)
    top = (block["J"] == NY - 1) & (block["I"] > 5) & (block["I"] < NX - 6)
    assert np.all(np.abs(lam[top] - 0.5) < 0.05)


def test_normals_point_out_of_each_face(block):
    n, I, J = block["normal"], block["I"], block["J"]
    mid_x, mid_y = (I > 5) & (I < NX - 6), (J > 5) & (J < NY - 6)
    for face, direction in (((J == NY - 1) & mid_x, (0, 1)), ((J == 0) & mid_x, (0, -1)),
                            ((I == 0) & mid_y, (-1, 0)), ((I == NX - 1) & mid_y, (1, 0))):
        assert face.sum() > 10
        assert np.all(n[face] @ np.array(direction, float) > 0.999)


def test_surface_classes_on_block(block):
    surf, d = block["surf"], block["d_edge"]
    h = block["h"]
    assert set(np.unique(surf)) <= {0.0, 2.0 / 3.0, 1.0, 2.0}
    # the outermost layer fails the angle test; rows within one support of
    # it take the projected class; rows beyond never see a masked neighbour
    assert np.all(surf[d < DX] == 0)
    assert np.all(surf[(d > DX) & (d < h)] == 1)
    assert np.all(surf[d > h + DX] == 2)


def test_shift_vanishes_on_lattice_and_restores_displacement(block):
    f, phys, h = block["f"], block["phys"], block["h"]
    raw, _ = O.dplus_shift(f, phys, DX / h, block["ff"], block["normal"], block["surf"])
    bulk = block["d_edge"] > h + DX
    scale = np.abs(raw).max()
    assert np.abs(raw[bulk]).max() < 1e-4 * scale
    k = (NY // 2) * NX + NX // 2
    x2 = f["x"].copy()
    x2[k, 0] += 0.1 * DX
    ff2 = O.brute(x2.astype(np.float32), block["hh"], x2.astype(np.float32))
    raw2, _ = O.dplus_shift(dict(f, x=x2), phys, DX / h, ff2, block["normal"], block["surf"])
    assert raw2[k, 0] < 0 and abs(raw2[k, 1]) < 1e-6 * abs(raw2[k, 0])
    assert raw2[k + 1, 0] > 0      # the crowded neighbour moves away


def test_shift_is_proportional_to_speed(block):
    """Ma = max|v| / c0 scales the whole shift linearly (before the clamp)."""
    f, phys, h = block["f"], block["phys"], block["h"]
    r1, _ = O.dplus_shift(f, phys, DX / h, block["ff"], block["normal"], block["surf"])
    r3, _ = O.dplus_shift(dict(f, v=3 * f["v"]), phys, DX / h, block["ff"], block["normal"],
                          block["surf"])
    assert np.allclose(r3, 3 * r1, rtol=1e-12, atol=0)


def test_shift_adjustment(block):
    f, phys, h = block["f"], block["phys"], block["h"]
    rng = np.random.default_rng(0)
    x2 = f["x"] + rng.uniform(-0.2, 0.2, f["x"].shape) * DX
    ff2 = O.brute(x2.astype(np.float32), block["hh"], x2.astype(np.float32))
    f2 = dict(f, x=x2, v=np.tile([20.0, 0.0], (len(x2), 1)))
    raw, sh = O.dplus_shift(f2, phys, DX / h, ff2, block["normal"], block["surf"])
    surf, n = block["surf"], block["normal"]
    lens = np.linalg.norm(sh, axis=1)
    assert np.all(lens <= 0.05 * h * (1 + 1e-12))
    assert (np.linalg.norm(raw, axis=1) > 0.05 * h).any()     # the clamp acted somewhere
    assert np.all(sh[surf == 0] == 0)
    bulk = surf == 2
    capped = np.minimum(1, 0.05 * h / np.maximum(np.linalg.norm(raw, axis=1), 1e-300))
    assert np.allclose(sh[bulk], raw[bulk] * capped[bulk, None], rtol=1e-12, atol=1e-18)
    proj = (surf == 1) & (np.linalg.norm(n, axis=1) > 0.99)
    assert proj.sum() > 20
    along = np.abs(np.einsum("nd,nd->n", sh[proj], n[proj]))
    assert np.all(along <= 0.03 * np.maximum(lens[proj], 1e-30) + 1e-18)


def test_dplus_pressure_reduces_to_plain_pressure(block):
    f, phys, ff = block["f"], block["phys"], block["ff"]
    rng = np.random.default_rng(1)
    n = len(f["x"])
    # p_i >= 0 everywhere: the switch always takes p_i + p_j
    f_pos = dict(f, rho=1 + rng.uniform(0, 0.01, n))
    a = O.dplus_pressure(f_pos, phys, ff, block["surf"], terms=O.T_PRESS)
    _, ref, _ = O.dsph_rhs(f_pos, phys, ff, None, None, O.T_PRESS)
    assert np.allclose(a, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # negative pressure, surface classes (< 1.5): also p_i + p_j
    f_neg = dict(f, rho=1 - rng.uniform(0, 0.01, n))
    a = O.dplus_pressure(f_neg, phys, ff, np.ones(n), terms=O.T_PRESS)
    _, ref, _ = O.dsph_rhs(f_neg, phys, ff, None, None, O.T_PRESS)
    assert np.allclose(a, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # uniform negative pressure in the bulk class: p_j - p_i = 0
    f_uni = dict(f, rho=np.full(n, 0.995))
    a = O.dplus_pressure(f_uni, phys, ff, np.full(n, 2.0), terms=O.T_PRESS)
    assert np.all(a == 0)


def test_boundary_rows_split_lists_equal_whole_lattice(block):
    """norm_rows over two lists (a split of each row's neighbours into
    'fluid' and 'boundary' points) gives the whole lattice's M, L and eig."""
    x, f, ff, h = block["x"], block["f"], block["ff"], block["h"]
    is_b = x[:, 0] < 0.1
    bidx, fidx = np.nonzero(is_b)[0], np.nonzero(~is_b)[0]
    xb, xf = x[bidx], x[fidx]
    hb = np.full(len(xb), h, np.float32)
    bf = O.brute(xb, hb, xf)
    bb = O.brute(xb, hb, xb)
    vf = f["area"][fidx] / f["rho"][fidx]
    vb = f["area"][bidx] / f["rho"][bidx]
    M, Lm, e = O.norm_rows(xb, h, bf, xf, vf, bb, xb, vb)
    assert np.allclose(M, block["M"][bidx], atol=1e-12)
    assert np.allclose(Lm, block["L"][bidx], atol=1e-9)
    assert np.allclose(e, block["eig"][bidx], atol=1e-12)
