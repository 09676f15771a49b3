"""Synthetic scenes for the hot path (dam breaks, colliding drops, point clouds).

Scene initialisation is outside the hot path; these generators reproduce the
reference's geometric contract rather than its init code:
  - fluid on a lattice x = (i + 1/2) dx with area A = dx^2, radius
    r = sqrt(A / pi) and support h = sqrt(A / pi * targetNeighbors) in fp32
    (simulationBase.py:854-855);
  - solid (mDBC) walls: `layers` rows of boundary particles outside each wall
    at (k + 1/2) dx with ghosts mirrored across the wall
    (solidBoundary.py:139-161, ghostParticles.py:128-133), boundary volume
    pi r^2 (solidBoundary.py:196-197);
  - Akinci walls: one row outside the fluid, volume gamma / sum_b' W_bb' and
    density term sum_b' W_bb' V_b' (akinciBoundary.py:137-157), over the wall
    pairs of this package's own host search (pairs_within, the fp32
    predicate |x_b - x_b'| < h) instead of torch_cluster's radius();
  - search domain = tank +/- 2 h (simulationBase.py:968-969).
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81


@dataclass
class Scene:
    name: str
    scheme: str                 # "dsph" (delta-SPH, RK4) or "dfsph" (semi-implicit)
    boundary: str               # "solid" (mDBC) or "akinci"
    x: np.ndarray               # [n, 2] fp32
    v: np.ndarray               # [n, 2] fp32
    area: float
    support: float              # h (fp32 value)
    rho0: float
    bx: np.ndarray              # [nb, 2] fp32 boundary positions
    bgx: np.ndarray | None      # [nb, 2] ghost positions (solid)
    bvol: np.ndarray            # [nb] boundary volume
    bterm: np.ndarray | None    # [nb] akinci density term
    vmin: tuple                 # search domain
    vmax: tuple
    phys: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.x)

    @property
    def nb(self) -> int:
        return len(self.bx)


def support_for(dx: float, target_neighbors: int = 20) -> np.float32:
    area = dx * dx
    return np.float32(math.sqrt(area / math.pi * target_neighbors))


# Rest-state calibration of the lattice set-ups, the counterpart of the
# reference's packing and boundary-spacing optimisation
# (simulationBase.py:857-880, evalPacking / evalBoundarySpacing): there the
# lattice spacing and the fluid-to-wall gap are solved so that a particle of
# the initial lattice has SPH density 1.  Here the spacing dx is given (it
# sets the particle count), so the particle area is scaled instead, and the
# Akinci wall layer is placed at the distance that gives the first fluid row
# density 1.  Both are ratios independent of dx (h ~ dx, areas ~ dx^2).
AKINCI_GAMMA = 0.7


@functools.lru_cache(maxsize=None)
def lattice_calibration(target_neighbors: int = 20, gamma: float = AKINCI_GAMMA):
    """(area / dx^2, Akinci wall offset / dx): the area that gives a square
    lattice of spacing dx density sum_j A W_ij = 1, and the distance below the
    fluid's edge (first row at dx / 2) of a wall line of spacing dx, with
    volumes gamma / sum W over the line, that gives the first row density 1."""
    dx = 1.0
    h = float(support_for(dx, target_neighbors))
    r = int(math.ceil(h / dx)) + 2

    def w(d):
        q = np.minimum(d / h, 1.0)
        return np.where(d < h, (1.0 - q) ** 4 * (1.0 + 4.0 * q) * 7.0 / (math.pi * h * h), 0.0)

    I, J = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="xy")
    lat = np.stack([I.ravel(), J.ravel()], 1) * dx
    area = 1.0 / w(np.sqrt((lat ** 2).sum(1))).sum()
    xs = np.arange(-r, r + 1) * dx
    vb = gamma / w(np.abs(xs)).sum()
    half = lat[lat[:, 1] >= 0] + np.array([0.0, 0.5 * dx])    # rows dx/2, 3dx/2, ...
    q = np.array([0.0, 0.5 * dx])
    rho_fluid = area * w(np.sqrt(((half - q) ** 2).sum(1))).sum()

    def rho_first(off):
        wall = np.stack([xs, np.full_like(xs, -off)], 1)
        return rho_fluid + vb * w(np.sqrt(((wall - q) ** 2).sum(1))).sum()

    lo, hi = 0.0, 2.0 * dx     # rho_first decreases with the offset
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if rho_first(mid) > 1.0:
            lo = mid
        else:
            hi = mid
    return area / (dx * dx), 0.5 * (lo + hi) / dx


# DFSPH time steps as dt / h (s per m of support), the invariant that carries a
# config's step to another resolution at the same geometry and speeds.
#  - colliding drops: configs/collision_dfsph.toml steps its fixed
#    integration.dt = 0.0025 s ([timestep] fixed = true keeps
#    simulationState['dt'] = integration.dt, simulationBase.py:433,
#    adaptiveDT.py:61-62) at radius 0.025 and 30 target neighbours, support
#    0.025 sqrt(30) (simulationBase.py:854-855);
#  - dam break: configs/dambreak_dfsph.toml steps integration.dt = 0.005 s at
#    radius 0.01, support 0.01 sqrt(20), i.e. 0.112 s/m.  That step holds the
#    reference's solver (no warm start, mean-residual stop) only on columns
#    of about 50 particle rows, the size of the shipped scene: with the FP64
#    oracle, 0.112 s/m diverges at 140 and more rows, 0.03 s/m at 316 rows.
#    The dam breaks therefore step the config's [timestep] bound instead,
#    0.0004 s at the same support (min = max = 0.0004), a chosen value that
#    is not the step the reference takes with fixed = true.
DFSPH_DT_PER_H_DROPS = 0.0025 / float(np.float32(0.025 * math.sqrt(30.0)))
DFSPH_DT_PER_H_DAMBREAK = 0.0004 / float(np.float32(0.01 * math.sqrt(20.0)))
# delta-SPH colliding drops: configs/collision_deltasph.toml sets fluid.c0 =
# 100, kinematic viscosity 0 and a fixed integration.dt = 0.001 s at radius
# 0.025 and 50 target neighbours (acoustic CFL c0 dt / h = 0.57).  The drops
# meet at 4 m/s and throw jets of 10-25 m/s: with the dam-break sound speed
# (10 sqrt(2 g 0.3) = 24 m/s, momentum.py:62-66) those jets reach Mach 1 and
# go through the floor's mDBC layers.
DSPH_C0_DROPS = 100.0
DSPH_DT_PER_H_DROPS = 0.001 / float(np.float32(0.025 * math.sqrt(50.0)))


def default_phys(scheme: str, h: float, rho0: float = 1000.0, kind: str = "dambreak") -> dict:
    """Solver constants of the shipped configs (configs/dambreak_deltasph.toml,
    configs/dambreak_dfsph.toml, configs/collision_dfsph.toml); `kind` picks
    the DFSPH time step of the scene family."""
    if scheme == "dsph":
        if kind == "drops":
            c0, dt, nu = DSPH_C0_DROPS, DSPH_DT_PER_H_DROPS * h, 0.0
        else:
            c0 = 10.0 * math.sqrt(2.0 * G * 0.3)   # momentum.py:62-66
            dt = 0.4 * h / c0                       # CFL 0.4, fixed
            nu = 0.001
        return dict(h=h, rho0=rho0, c0=c0, gamma=7.0, kappa=1.3, alpha=0.01, delta=0.1,
                    nu=nu, eps=0.1 * h * h, gx=0.0, gy=-G, dt=dt, beta=0.125, omega=0.5,
                    tau=1e-3, fluid_omega=0.3, min_iters=2, max_iters=256, eos=0,
                    diffusion="MOG", boundary_diffusion=True)
    c0 = 100.0
    ratio = DFSPH_DT_PER_H_DROPS if kind == "drops" else DFSPH_DT_PER_H_DAMBREAK
    dt = ratio * h
    # [dfsph] of both shipped DFSPH configs; divergenceSolver = false there
    return dict(h=h, rho0=rho0, c0=c0, gamma=7.0, kappa=1.3, alpha=0.01, delta=0.1, nu=0.01,
                eps=0.1 * h * h, gx=0.0, gy=-G, dt=dt, beta=0.125, omega=0.5, tau=1e-3,
                fluid_omega=0.3, min_iters=2, max_iters=256, eos=0, diffusion="none",
                boundary_diffusion=False, divergence=False, div_tau=1e-2, div_min_iters=2,
                div_max_iters=8)


def dplus_variant(scene: Scene, shifting: bool = True) -> Scene:
    """The delta-SPH scene under the delta-plus scheme: surface detection
    and the pressure switch in every right-hand side, and with `shifting`
    the particle shift after each step (deltaSPH.py:119-122,
    simulationBase.py:708-732)."""
    import dataclasses

    assert scene.scheme == "dsph" and scene.boundary == "solid"
    phys = dict(scene.phys, dplus=True, shifting=bool(shifting))
    return dataclasses.replace(scene, name=scene.name + "_dplus", phys=phys)


def _jitter(rng, pts, dx, amp=0.01):
    return pts + rng.uniform(-amp * dx, amp * dx, size=pts.shape)


def box_normals(bx, lo, hi) -> np.ndarray:
    """Outward unit normals of wall points outside the box [lo, hi]: the
    gradient of the box's signed distance, (p - nearest box point) / |.|
    (the polygon SDF gradient of solidBoundary.py:158 for a rectangle)."""
    bx = np.asarray(bx, dtype=np.float64)
    d = bx - np.clip(bx, np.asarray(lo, np.float64), np.asarray(hi, np.float64))
    n = np.linalg.norm(d, axis=1, keepdims=True)
    return (d / np.where(n > 0, n, 1.0)).astype(np.float32)


def _walls_solid(W, H, dx, layers=3):
    nx, ny = int(round(W / dx)), int(round(H / dx))
    i = np.arange(-layers, nx + layers)
    j = np.arange(-layers, ny + layers)
    I, J = np.meshgrid(i, j, indexing="xy")
    I, J = I.ravel(), J.ravel()
    outside = (I < 0) | (I >= nx) | (J < 0) | (J >= ny)
    bx = np.stack([(I[outside] + 0.5) * dx, (J[outside] + 0.5) * dx], axis=1)
    gx = bx.copy()
    lo = bx[:, 0] < 0
    hi = bx[:, 0] > W
    gx[lo, 0] = -bx[lo, 0]
    gx[hi, 0] = 2 * W - bx[hi, 0]
    lo = bx[:, 1] < 0
    hi = bx[:, 1] > H
    gx[lo, 1] = -bx[lo, 1]
    gx[hi, 1] = 2 * H - bx[hi, 1]
    return bx, gx


def _walls_akinci(W, H, dx):
    nx, ny = int(round(W / dx)), int(round(H / dx))
    off = lattice_calibration()[1] * dx
    xs = (np.arange(-1, nx + 1) + 0.5) * dx
    ys = (np.arange(0, ny) + 0.5) * dx
    bottom = np.stack([xs, np.full_like(xs, -off)], 1)
    top = np.stack([xs, np.full_like(xs, H + off)], 1)
    left = np.stack([np.full_like(ys, -off), ys], 1)
    right = np.stack([np.full_like(ys, W + off), ys], 1)
    return np.concatenate([bottom, top, left, right])


def _wendland_np(q, h):
    a = 1.0 - q
    return a ** 4 * (1.0 + 4.0 * q) * (7.0 / math.pi) / (h * h)


def pairs_within(x: np.ndarray, h: float) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """All ordered pairs (i, j), self pairs included, with fp32
    |x_i - x_j| < h (the neighbour predicate of the search,
    neighSearch.cpp:93-107), by a uniform grid of cell size h on the host:
    points are sorted by cell, and each point's 3x3 cell stencil is a set of
    contiguous runs of the sorted order.  The init-time counterpart of the
    reference's torch_cluster radius() call (akinciBoundary.py:137,
    solidBoundary.py:180).  Returns (i, j, fp32 distance)."""
    xf = np.ascontiguousarray(x, dtype=np.float32)
    n = len(xf)
    if n == 0:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros(0, np.float32)
    h32 = np.float32(h)
    c = np.floor((xf - xf.min(0)) / h32).astype(np.int64)
    ny = int(c[:, 1].max()) + 3
    key = (c[:, 0] + 1) * ny + (c[:, 1] + 1)
    order = np.argsort(key, kind="stable")
    skey = key[order]
    ii, jj = [], []
    for ox in (-1, 0, 1):
        for oy in (-1, 0, 1):
            t = key + ox * ny + oy
            lo = np.searchsorted(skey, t, side="left")
            hi = np.searchsorted(skey, t, side="right")
            cnt = hi - lo
            tot = int(cnt.sum())
            if tot == 0:
                continue
            start = np.repeat(np.cumsum(cnt) - cnt, cnt)
            ii.append(np.repeat(np.arange(n), cnt))
            jj.append(order[np.arange(tot) - start + np.repeat(lo, cnt)])
    i, j = np.concatenate(ii), np.concatenate(jj)
    d = np.sqrt(((xf[i] - xf[j]) ** 2).sum(1, dtype=np.float32)).astype(np.float32)
    keep = d < h32
    return i[keep], j[keep], d[keep]


def akinci_volumes(bx: np.ndarray, h: float, gamma: float = 0.7):
    """V_b = gamma / sum_b' W_bb' and term_b = sum_b' W_bb' V_b'
    (akinciBoundary.py:152-157), boundary pairs |x_b - x_b'| < h from
    pairs_within."""
    i, j, d = pairs_within(bx, h)
    d = d.astype(np.float64)
    w = _wendland_np(d / h, h)
    wsum = np.bincount(i, weights=w, minlength=len(bx))
    vol = gamma / wsum
    term = np.bincount(i, weights=w * vol[j], minlength=len(bx))
    return vol.astype(np.float32), term.astype(np.float32)


def _finish(name, scheme, boundary, fluid, vel, dx, W, H, seed, h=None, kind="dambreak"):
    h = support_for(dx) if h is None else np.float32(h)
    area = np.float32(dx * dx * lattice_calibration()[0])
    if boundary == "solid":
        bx, bgx = _walls_solid(W, H, dx)
        bvol = np.full(len(bx), area, dtype=np.float32)   # pi r^2 = A
        bterm = None
    else:
        bx = _walls_akinci(W, H, dx)
        bgx = None
        bvol, bterm = akinci_volumes(bx, h, AKINCI_GAMMA)
    phys = default_phys(scheme, float(h), kind=kind)
    buf = 2.0 * float(h)
    return Scene(name=name, scheme=scheme, boundary=boundary, x=fluid.astype(np.float32),
                 v=vel.astype(np.float32), area=float(area), support=float(h),
                 rho0=1000.0, bx=bx.astype(np.float32),
                 bgx=None if bgx is None else bgx.astype(np.float32), bvol=bvol, bterm=bterm,
                 vmin=(-buf, -buf), vmax=(W + buf, H + buf), phys=phys)


def dam_break(dx=0.01, block=(1.0, 0.5), tank=(4.0, 2.0), scheme="dsph", boundary=None,
              seed=0) -> Scene:
    """Fluid block [0, bw] x [0, bh] in a closed tank; jitter U(+-0.01 dx)."""
    boundary = boundary or ("solid" if scheme == "dsph" else "akinci")
    rng = np.random.default_rng(seed)
    nx, ny = int(round(block[0] / dx)), int(round(block[1] / dx))
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    pts = np.stack([(I.ravel() + 0.5) * dx, (J.ravel() + 0.5) * dx], axis=1)
    pts = _jitter(rng, pts, dx)
    vel = np.zeros_like(pts)
    return _finish(f"dambreak_{scheme}_{len(pts)}", scheme, boundary, pts, vel, dx, *tank, seed)


def colliding_drops(dx=0.01, radius=0.25, speed=2.0, tank=(4.0, 2.0), scheme="dsph",
                    boundary=None, seed=0) -> Scene:
    """Two discs approaching each other at +-speed (configs/collision_*.toml)."""
    boundary = boundary or ("solid" if scheme == "dsph" else "akinci")
    rng = np.random.default_rng(seed)
    W, H = tank
    pts, vel = [], []
    for cx, u in ((W / 2 - 2 * radius, speed), (W / 2 + 2 * radius, -speed)):
        n = int(math.ceil(radius / dx)) + 1
        I, J = np.meshgrid(np.arange(-n, n), np.arange(-n, n), indexing="xy")
        p = np.stack([(I.ravel() + 0.5) * dx, (J.ravel() + 0.5) * dx], axis=1)
        p = p[(p ** 2).sum(1) <= radius * radius]
        p = p + np.array([cx, H / 2])
        pts.append(p)
        vel.append(np.tile([u, 0.0], (len(p), 1)))
    pts = _jitter(rng, np.concatenate(pts), dx)
    vel = np.concatenate(vel)
    return _finish(f"drops_{scheme}_{len(pts)}", scheme, boundary, pts, vel, dx, W, H, seed,
                   kind="drops")


def scene_for_count(kind: str, scheme: str, n_target: int, seed: int) -> Scene:
    """Dam break or colliding drops with dx chosen to give about n_target
    fluid particles."""
    if kind == "dambreak":
        dx = math.sqrt(1.0 * 0.5 / n_target)
        return dam_break(dx=dx, scheme=scheme, seed=seed)
    area = 2 * math.pi * 0.25 ** 2
    dx = math.sqrt(area / n_target)
    return colliding_drops(dx=dx, scheme=scheme, seed=seed)


def point_cloud(n: int, seed: int = 1, clusters: int = 8, sigma: float = 0.05) -> np.ndarray:
    """Half uniform in [0,1)^2, half from Gaussian clusters (sigma 0.05,
    centres U[0.15, 0.85]^2), resampling points that leave the box."""
    rng = np.random.default_rng(seed)
    nu = n // 2
    pts = [rng.uniform(0.0, 1.0, size=(nu, 2))]
    centres = rng.uniform(0.15, 0.85, size=(clusters, 2))
    per = [(n - nu) // clusters + (1 if k < (n - nu) % clusters else 0) for k in range(clusters)]
    for c, m in zip(centres, per):
        out = np.empty((0, 2))
        while len(out) < m:
            p = rng.normal(c, sigma, size=(2 * (m - len(out)) + 16, 2))
            p = p[(p >= 0).all(1) & (p < 1).all(1)]
            out = np.concatenate([out, p])
        pts.append(out[:m])
    return np.concatenate(pts).astype(np.float32)


def support_for_mean_neighbors(pts: np.ndarray, target: float = 20.0, sample: int = 20000,
                               seed: int = 0) -> np.float32:
    """Support radius giving `target` mean neighbours (self included),
    bisected on a sample of query points against all points."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    tree = cKDTree(pts.astype(np.float64))
    qs = pts[rng.choice(len(pts), size=min(sample, len(pts)), replace=False)].astype(np.float64)
    # bracket around the radius of a uniform cloud of the same size
    r0 = float(np.sqrt(target / (np.pi * len(pts))))
    lo, hi = r0 / 8, r0 * 8
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        cnt = float(np.mean(tree.query_ball_point(qs, mid, return_length=True)))
        if cnt < target:
            lo = mid
        else:
            hi = mid
    return np.float32(0.5 * (lo + hi))
