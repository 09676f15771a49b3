"""GPU neighbour search: pair sets bit-exact against the oracle's restatement
of the reference grid search (neighSearch.cpp:172-331)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pytorchsph_amd import scenes as S
from pytorchsph_amd import _lib as L
from pytorchsph_amd.search import NeighborGrid, grid_params, pair_geometry, transpose
from tests.gpu_helpers import assert_same_pairs, gpu_search, sorted_rows

pytestmark = pytest.mark.gpu


def oracle_self(x, h, vmin, vmax, hmax):
    qx, qy, cx, _, _ = grid_params(vmin, vmax, hmax)
    return O.search(x, h, x, (qx, qy), hmax, cx, 1, self_mode=True)


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 1), (2000, 2), (50000, 3)])
def test_uniform_cloud(cuda, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    hmax = np.float32(np.sqrt(20 / (np.pi * max(n, 1))))
    h = np.full(n, hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert st == 0
    assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))


def test_clustered_cloud(cuda):
    x = S.point_cloud(200_000, seed=1)
    hmax = S.support_for_mean_neighbors(x, 20.0, sample=5000)
    h = np.full(len(x), hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert st == 0
    assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))
    mean = (rp[-1]) / len(x)
    assert 15 < mean < 25


def test_variable_support(cuda):
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 1, (20000, 2)).astype(np.float32)
    hmax = np.float32(0.02)
    h = (rng.uniform(0.3, 1.0, len(x)) * hmax).astype(np.float32)
    h[0] = hmax
    rp, c, _, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))


def test_adversarial_points(cuda):
    """Near-tie distances (+-3 ulp of h), points exactly on cell faces,
    duplicates, a dense single cell and points at the domain corner."""
    hmax = np.float32(0.0625)
    qx, qy, cx, cy, _ = grid_params((0, 0), (1, 1), hmax)
    pts = [[0.5, 0.5]]
    for k in range(-3, 4):
        d = hmax
        for _ in range(abs(k)):
            d = np.nextafter(d, np.float32(1 if k > 0 else 0))
        pts += [[0.5 + d, 0.5], [0.5, 0.5 - d], [0.5 - d, 0.5]]
    faces = qx + hmax * np.arange(2, 12, dtype=np.float32)
    pts += [[f, 0.3] for f in faces] + [[0.3, f] for f in faces]
    pts += [[0.25, 0.25]] * 5                                   # duplicates
    rng = np.random.default_rng(1)
    pts += list(0.8 + rng.uniform(0, 0.01, (60, 2)))            # one dense cell
    pts += [[0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0]]
    x = np.array(pts, np.float32)
    h = np.full(len(x), hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert st == 0
    assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))


def test_foreign_queries_and_transpose(cuda):
    sc = S.dam_break(dx=0.01, scheme="dsph", seed=0)
    hmax = np.float32(sc.support)
    qx, qy, cx, _, _ = grid_params(sc.vmin, sc.vmax, hmax)
    bh = np.full(sc.nb, hmax, np.float32)
    for q in (sc.bx, sc.bgx):
        rp, c, st, g = gpu_search(sc.x, np.full(sc.n, hmax, np.float32), sc.vmin, sc.vmax, hmax,
                                  queries=(q, bh))
        orc = O.search(q, bh, sc.x, (qx, qy), hmax, cx, 1)
        assert_same_pairs(rp, c, orc)
    # device transpose of boundary->fluid equals the oracle transpose
    dev = torch.device("cuda")
    from pytorchsph_amd.search import CSR
    a = CSR(torch.from_numpy(rp.astype(np.int32)).to(dev),
            torch.from_numpy(c.astype(np.int32)).to(dev), len(rp) - 1, int(rp[-1]))
    t = transpose(a, sc.n)
    trp = t.rowptr.cpu().numpy().astype(np.int64)
    tcol = t.col[: int(trp[-1])].cpu().numpy().astype(np.int64)
    ot = O.transpose(orc, sc.n)
    assert np.array_equal(trp, ot.rowptr) and np.array_equal(tcol, ot.col)


def test_multi_scene_grid(cuda):
    """Three scenes with different domains and supports in one grid; each
    scene's pairs equal its own single-scene oracle result."""
    rng = np.random.default_rng(4)
    specs = [((0, 0), (1, 1), 0.03, 3000), ((-2, -1), (2, 1), 0.05, 5000),
             ((10, 10), (10.5, 12), 0.02, 4000)]
    xs, scenes, off = [], [], 0
    for vmin, vmax, h, n in specs:
        x = rng.uniform(vmin, vmax, (n, 2)).astype(np.float32)
        xs.append(x)
        scenes.append(dict(vmin=vmin, vmax=vmax, hmax=np.float32(h), p_begin=off, p_count=n))
        off += n
    X = np.concatenate(xs)
    H = np.concatenate([np.full(n, np.float32(h), np.float32) for _, _, h, n in specs])
    P = np.concatenate([np.full(n, s, np.int32) for s, (_, _, _, n) in enumerate(specs)])
    dev = torch.device("cuda")
    g = NeighborGrid(scenes, len(X), dev)
    xt, ht, pt = (torch.from_numpy(a).to(dev) for a in (X, H, P))
    g.build(xt, pt)
    csr = g.search(xt, ht, pt, 1, self_sorted=True, sort_rows=True)
    rp = csr.rowptr.cpu().numpy().astype(np.int64)
    col = csr.col[: int(rp[-1])].cpu().numpy().astype(np.int64)
    off = 0
    for (vmin, vmax, h, n), x in zip(specs, xs):
        orc = oracle_self(x, np.full(n, np.float32(h), np.float32), vmin, vmax, np.float32(h))
        seg_rp = rp[off:off + n + 1] - rp[off]
        seg_col = col[rp[off]:rp[off + n]] - off
        assert np.array_equal(seg_rp, orc.rowptr)
        assert np.array_equal(seg_col, orc.col)
        off += n


def test_pair_geometry_matches_reference_layout(cuda):
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (3000, 2)).astype(np.float32)
    hmax = np.float32(0.05)
    h = np.full(len(x), hmax, np.float32)
    dev = torch.device("cuda")
    g = NeighborGrid([dict(vmin=(0, 0), vmax=(1, 1), hmax=hmax, p_begin=0, p_count=len(x))],
                     len(x), dev)
    xt, ht = torch.from_numpy(x).to(dev), torch.from_numpy(h).to(dev)
    ps = torch.zeros(len(x), dtype=torch.int32, device=dev)
    g.build(xt, ps)
    csr = g.search(xt, ht, ps, 1, self_sorted=True, sort_rows=True)
    rows, cols, rhat, q = pair_geometry(csr, xt, xt, None, float(hmax))
    rows, cols = rows.cpu().numpy(), cols.cpu().numpy()
    orc = oracle_self(x, h, (0, 0), (1, 1), hmax)
    r_o, c_o = orc.pairs()
    assert np.array_equal(rows, r_o) and np.array_equal(cols, c_o)   # lexicographic (i, j)
    d = x[rows] - x[cols]
    dist = np.linalg.norm(d.astype(np.float64), axis=1)
    want = np.where(dist[:, None] >= 1e-7, d / np.maximum(dist, 1e-30)[:, None], 0.0)
    assert np.allclose(rhat.cpu().numpy(), want, atol=2e-6)
    assert np.allclose(q.cpu().numpy(), dist / hmax, atol=2e-6)
    assert np.all(rhat.cpu().numpy()[rows == cols] == 0.0)


@pytest.mark.parametrize("n", [0, 1])
def test_empty_and_single_point(cuda, n):
    """An empty point set gives an empty list; one point pairs with itself
    only (the reference's self pair)."""
    x = np.full((n, 2), 0.5, np.float32)
    hmax = np.float32(0.05)
    rp, c, st, _ = gpu_search(x, np.full(n, hmax, np.float32), (0, 0), (1, 1), hmax)
    assert st == 0
    assert rp.tolist() == list(range(n + 1))
    assert c.tolist() == list(range(n))


def test_walk_count_crowded_chunks_and_long_rows(cuda):
    """The LDS-staged count where it cannot stage and where rows are long:
    3,000 points in a tenth of a cell (every one a neighbour of all; their
    chunks' stencils are far larger than the LDS stage and are counted from
    global memory) and 600 points in a patch of 1.5 x 1.5 supports (rows of
    ~300 hits, staged) beside 20,000 uniform points.  Pairs bit-exact
    against the oracle (the fill writes exactly the counted rows)."""
    rng = np.random.default_rng(12)
    n_u = 20000
    hmax = np.float32(np.sqrt(20 / (np.pi * n_u)))
    x = np.concatenate([rng.uniform(0, 1, (n_u, 2)),
                        0.5 + rng.uniform(0, 0.1 * hmax, (3000, 2)),
                        0.25 + rng.uniform(0, 1.5 * hmax, (600, 2))]).astype(np.float32)
    h = np.full(len(x), hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert st == 0
    rows = np.diff(rp)
    assert rows.max() >= 3000 and (rows > 200).sum() >= 500
    assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))


def test_overflow_is_flagged_and_bounded(cuda):
    rng = np.random.default_rng(8)
    x = rng.uniform(0, 1, (4000, 2)).astype(np.float32)
    hmax = np.float32(0.05)
    h = np.full(len(x), hmax, np.float32)
    dev = torch.device("cuda")
    g = NeighborGrid([dict(vmin=(0, 0), vmax=(1, 1), hmax=hmax, p_begin=0, p_count=len(x))],
                     len(x), dev)
    xt, ht = torch.from_numpy(x).to(dev), torch.from_numpy(h).to(dev)
    ps = torch.zeros(len(x), dtype=torch.int32, device=dev)
    g.build(xt, ps)
    q = g.query_struct(xt, ht, ps, 1, True)
    counts = torch.zeros(len(x), dtype=torch.int32, device=dev)
    rowptr = torch.zeros(len(x) + 1, dtype=torch.int32, device=dev)
    g.count(q, counts)
    from pytorchsph_amd.search import scan
    scan(counts, len(x), rowptr)
    cap = int(rowptr[-1].item()) // 2
    col = torch.full((cap,), -1, dtype=torch.int32, device=dev)
    g.fill(q, rowptr, col)
    torch.cuda.synchronize()
    assert int(g.status[0].item()) & L.ST_OVERFLOW
    rp = rowptr.cpu().numpy()
    assert rp.max() <= cap and np.all(np.diff(rp) >= 0)
    assert int((col < 0).sum()) == 0


def test_points_outside_table_are_flagged(cuda):
    x = np.array([[0.5, 0.5], [0.51, 0.5], [50.0, 50.0]], np.float32)
    hmax = np.float32(0.05)
    rp, c, st, _ = gpu_search(x, np.full(3, hmax, np.float32), (0, 0), (1, 1), hmax)
    assert st & L.ST_OUTSIDE
    orc = oracle_self(x[:2], np.full(2, hmax, np.float32), (0, 0), (1, 1), hmax)
    assert np.array_equal(sorted_rows(rp[:3], c[: rp[2]]), orc.col.astype(np.int64))


def test_point_past_a_side_edge_is_flagged_and_keeps_reference_pairs(cuda):
    """A point past the right edge of the table aliases into a cell of the
    next row, as the reference's linear key does (neighSearch.cpp:96): its
    pairs stay the reference's, and the scene is flagged SPH_ST_OUTSIDE."""
    hmax = np.float32(0.05)
    qx, qy, cx, cy, _ = grid_params((0, 0), (1, 1), hmax)
    rng = np.random.default_rng(5)
    inside = rng.uniform(0.0, 1.0, (400, 2)).astype(np.float32)
    past = np.array([[qx + (cx + 0.5) * hmax, 0.5]], np.float32)   # ix == cells_x
    x = np.concatenate([inside, past])
    h = np.full(len(x), hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert st & L.ST_OUTSIDE
    assert_same_pairs(rp, c, oracle_self(x, h, (0, 0), (1, 1), hmax))
    # the inside points alone raise nothing
    _, _, st0, _ = gpu_search(inside, h[:-1], (0, 0), (1, 1), hmax)
    assert st0 == 0


@pytest.mark.slow
def test_c2_four_million_points(cuda):
    """Neighbour search alone at 4M points (half uniform, half 8 Gaussian
    clusters, ~20 mean neighbours), bit-exact against the oracle, with the
    grid of the radix-sort build (the benchmark's build for a point cloud)."""
    x = S.point_cloud(4_000_000, seed=1)
    hmax = S.support_for_mean_neighbors(x, 20.0)
    h = np.full(len(x), hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax, ordered=False)
    assert st == 0
    orc = oracle_self(x, h, (0, 0), (1, 1), hmax)
    assert_same_pairs(rp, c, orc)


GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "search_ref.npz"


@pytest.mark.parametrize("case", ["uniform", "clustered", "variable", "neartie"])
def test_device_matches_reference_fixture(cuda, case):
    """Pairs equal to those of the reference's own compiled CPU search."""
    g = np.load(GOLD)
    x, h = g[f"{case}_x"], g[f"{case}_h"]
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), np.float32(h.max()))
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    o = np.lexsort((c, rows))
    assert np.array_equal(rows[o], g[f"{case}_rows"]) and np.array_equal(c[o], g[f"{case}_cols"])


def test_device_foreign_matches_reference_fixture(cuda):
    g = np.load(GOLD)
    x, h = g["uniform_x"], g["uniform_h"]
    rp, c, _, _ = gpu_search(x, h, (0, 0), (1, 1), np.float32(h.max()),
                             queries=(g["foreign_xq"], g["foreign_hq"]))
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    o = np.lexsort((c, rows))
    assert np.array_equal(rows[o], g["foreign_rows"]) and np.array_equal(c[o], g["foreign_cols"])


@pytest.mark.parametrize("n", [0, 1, 7, 2047, 2048, 2049, 100_003, 5_000_000])
@pytest.mark.parametrize("offset", [0, 1])
def test_scan_matches_numpy(cuda, n, offset):
    """Exclusive prefix sum with the total at out[n], for aligned arrays (the
    int4 path) and arrays one int off alignment (the scalar path)."""
    from pytorchsph_amd.search import scan

    rng = np.random.default_rng(n + offset)
    vals = rng.integers(0, 60, size=n).astype(np.int32)
    dev = torch.device("cuda")
    inb = torch.zeros(n + offset + 1, dtype=torch.int32, device=dev)
    outb = torch.full((n + offset + 2,), -7, dtype=torch.int32, device=dev)
    inp = inb[offset:offset + n]
    inp.copy_(torch.from_numpy(vals))
    out = outb[offset:offset + n + 1]
    scan(inp, n, out)
    torch.cuda.synchronize()
    want = np.zeros(n + 1, dtype=np.int64)
    want[1:] = np.cumsum(vals)
    assert np.array_equal(out.cpu().numpy(), want.astype(np.int32))
    rest = outb.cpu().numpy()
    assert (rest[:offset] == -7).all() and (rest[offset + n + 1:] == -7).all()


def test_non_finite_points_have_no_neighbours(cuda):
    """NaN / Inf positions fail the fp32 predicate against every point, as in
    the reference: their rows are empty and they appear in no row.  They go
    to the scene's outside slot (SPH_ST_OUTSIDE) instead of piling into one
    cell, and the finite points' rows are the search without them."""
    rng = np.random.default_rng(21)
    x = rng.uniform(0.0, 1.0, size=(20_000, 2)).astype(np.float32)
    hmax = np.float32(0.02)
    bad = rng.choice(len(x), 3000, replace=False)
    x[bad[:1000], 0] = np.nan
    x[bad[1000:2000], 1] = np.inf
    x[bad[2000:], :] = -np.inf
    h = np.full(len(x), hmax, np.float32)
    rp, c, st, _ = gpu_search(x, h, (0, 0), (1, 1), hmax)
    assert st & L.ST_OUTSIDE
    cnt = np.diff(rp)
    assert (cnt[bad] == 0).all()
    assert not np.isin(c, bad).any()
    good = np.setdiff1d(np.arange(len(x)), bad)
    rp2, c2, st2, _ = gpu_search(x[good], h[good], (0, 0), (1, 1), hmax)
    s1 = sorted_rows(rp, c)
    s2 = sorted_rows(rp2, c2)
    for k, i in enumerate(good[:2000]):
        # good is ascending, so mapping the reduced indices keeps rows sorted
        assert np.array_equal(s1[rp[i]:rp[i + 1]], good[s2[rp2[k]:rp2[k + 1]]])


def exact_threshold(h):
    """Largest float32 T with fl(sqrt(T)) < h, from the FP64 midpoint:
    m = (pred(h) + h) / 2 has 25 significant bits, so m^2 is exact in FP64
    and is never a float32 (odd 49-50 bit significand)."""
    h = np.asarray(h, np.float32)
    hp = np.nextafter(h, np.float32(0))
    m = (hp.astype(np.float64) + h.astype(np.float64)) / 2
    m2 = m * m
    f = m2.astype(np.float32)
    over = f.astype(np.float64) >= m2
    f[over] = np.nextafter(f[over], np.float32(0))
    return f


def test_device_pair_thresholds_are_exact(cuda):
    """sph_pair_thresholds (the integer-arithmetic threshold every search
    query uses) against the FP64 definition, bit for bit: random supports,
    powers of two and their neighbours over 2^-60..2^60, and the supports
    of the benchmark scenes."""
    rng = np.random.default_rng(11)
    pow2 = np.ldexp(np.float32(1), np.arange(-60, 61)).astype(np.float32)
    bits = pow2.view(np.uint32)[:, None] + np.arange(-3, 4, dtype=np.int64)[None, :]
    h = np.concatenate([rng.uniform(1e-6, 1e3, 1_000_000).astype(np.float32),
                        bits.astype(np.uint32).ravel().view(np.float32),
                        np.array([S.support_for(d) for d in (0.0005, 0.001, 0.01, 0.02)],
                                 np.float32)])
    ht = torch.from_numpy(h).cuda()
    out = torch.empty_like(ht)
    L.call("sph_pair_thresholds", L.ptr(ht), len(h), L.ptr(out), L.stream_handle())
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), exact_threshold(h).view(np.uint32))


def _grid_arrays(g):
    torch.cuda.synchronize()
    return {k: getattr(g, k).cpu().numpy().copy()
            for k in ("sorted_idx", "sorted_x", "cell_start", "key", "status")}


@pytest.mark.parametrize("case", ["random", "clustered", "cell_order", "multi_scene",
                                  "outside_and_nan", "tiny", "keyed"])
def test_sorted_build_equals_atomic_build(cuda, case):
    """sph_grid_build_sorted (radix sort) gives sph_grid_build's grid bit for
    bit: the same sorted order (scene, key, index), positions, cell table,
    keys and status bits, on point clouds in random order, in cell order,
    over several scenes, with points outside the table and non-finite
    positions (fewer than the atomic build's unranked outside-slot limit),
    from one point up, and with the cell keys given."""
    rng = np.random.default_rng(11)
    dev = torch.device("cuda")
    keys = None
    if case in ("random", "clustered", "cell_order", "tiny", "keyed"):
        n = {"random": 300_000, "clustered": 300_000, "cell_order": 200_000, "tiny": 5,
             "keyed": 50_000}[case]
        x = (S.point_cloud(n, seed=3) if case == "clustered"
             else rng.uniform(0, 1, (n, 2))).astype(np.float32)
        hmax = np.float32(np.sqrt(20 / (np.pi * n)))
        scenes = [dict(vmin=(0, 0), vmax=(1, 1), hmax=hmax, p_begin=0, p_count=n)]
        P = np.zeros(n, np.int32)
        if case == "cell_order":
            qx, qy, cx, _, _ = grid_params((0, 0), (1, 1), hmax)
            k = (np.ceil((x[:, 0] - qx) / hmax) + cx * np.ceil((x[:, 1] - qy) / hmax))
            x = x[np.argsort(k, kind="stable")]
        if case == "keyed":
            qx, qy, cx, cy, _ = grid_params((0, 0), (1, 1), hmax)
            keys = (np.ceil((x[:, 0] - qx) / hmax) + cx * np.ceil((x[:, 1] - qy) / hmax))
            keys = keys.astype(np.int32)
            keys[::1000] = -5                    # out of range: the outside slot
    else:
        specs = [((0, 0), (1, 1), 0.03, 3000), ((-2, -1), (2, 1), 0.05, 5000),
                 ((10, 10), (10.5, 12), 0.02, 4000)]
        xs, scenes, off = [], [], 0
        for vmin, vmax, h, m in specs:
            xs.append(rng.uniform(vmin, vmax, (m, 2)).astype(np.float32))
            scenes.append(dict(vmin=vmin, vmax=vmax, hmax=np.float32(h), p_begin=off,
                               p_count=m))
            off += m
        x = np.concatenate(xs)
        P = np.concatenate([np.full(m, s, np.int32) for s, (_, _, _, m) in enumerate(specs)])
        if case == "outside_and_nan":
            x[5] = (50.0, 50.0)
            x[3001] = (-30.0, 0.0)
            x[3002] = (np.nan, 0.5)
            x[9000] = (10.25, np.inf)
            x[9001] = (10.9, 11.0)               # past the side edge: aliased key
    n = len(x)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    pt = torch.from_numpy(P).to(dev)
    out = []
    for ordered in (True, False):
        g = NeighborGrid(scenes, n, dev)
        g.x, g.pscene = xt, pt
        c = g._grid()
        if keys is None:
            g.build(xt, pt, ordered=ordered)
        else:
            kt = torch.from_numpy(keys).to(dev)
            if ordered:
                L.call("sph_grid_build_keyed", __import__("ctypes").byref(c), L.ptr(kt),
                       L.ptr(g.ws), g.ws_bytes, L.stream_handle())
            else:
                ws = g._sort_ws()
                L.call("sph_grid_build_sorted", __import__("ctypes").byref(c), L.ptr(kt),
                       L.ptr(ws), int(ws.numel()), L.stream_handle())
        out.append(_grid_arrays(g))
    a, b = out
    for k in a:
        if k == "sorted_x":
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (case, k)
        else:
            assert np.array_equal(a[k], b[k]), (case, k)
