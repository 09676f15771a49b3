"""Python face of the CPU oracle (sph_oracle.c).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg use this module,
always as the checker.  It restates the reference's algorithms:
  search          neighSearch.cpp:172-331, neighborSearch.py:30-116
  fluid_density   density.py:26-53, akinciBoundary.py:395-420
  dsph_rhs        deltaSPH.py:146-200 with momentum.py, densityDiffusion.py,
                  pressure.py, diffusion.py, laminar.py, gravity.py and the
                  solid boundary terms of solidBoundary.py:676-895
  mdbc            solidBoundary.py:561-616
  dplus_*         deltaPlus.py:81-251, pressure.py:40-52 (delta-plus shifting)
  dfsph           modules/dfsph.py:25-311 + akinciBoundary.py:166-389
  rk4 / semi-implicit steps   simulationBase.py:564-569, 632-705
Physics is FP64; the neighbour predicate is the reference's fp32 one.
Parity unpinned: the fp32 stopping sequence of densitySolve (see
sph_oracle.c); everything else is pinned by the tests named there.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

T_CONT, T_CONT_B, T_MOG, T_DDSPH = 1 << 0, 1 << 1, 1 << 2, 1 << 3
T_PRESS, T_PRESS_B, T_VISC, T_VISC_B = 1 << 4, 1 << 5, 1 << 6, 1 << 7
T_LAM, T_GRAV = 1 << 8, 1 << 9

_vp = C.c_void_p


class OrcCSR(C.Structure):
    _fields_ = [("nrows", C.c_int), ("rowptr", _vp), ("col", _vp)]


class OrcFluid(C.Structure):
    _fields_ = [("n", C.c_int), ("x", _vp), ("v", _vp), ("rho", _vp), ("area", _vp), ("rest", _vp),
                ("vol", _vp)]


class OrcBdy(C.Structure):
    _fields_ = [("n", C.c_int), ("x", _vp), ("v", _vp), ("rho", _vp), ("vol", _vp), ("p", _vp),
                ("rest", _vp), ("gx", _vp)]


class OrcPhys(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "h", "rho0", "c0", "gamma", "kappa", "alpha", "delta", "nu", "eps", "gx", "gy", "dt",
        "beta", "omega", "tau", "fluid_omega")] + [
        ("min_iters", C.c_int), ("max_iters", C.c_int), ("eos", C.c_int), ("pad", C.c_int)]


class OrcDfsphOut(C.Structure):
    _fields_ = [(n, _vp) for n in ("alpha", "src", "p2", "apred", "vpred", "b_alpha", "b_src",
                                   "b_p", "err_hist")] + [("max_hist", C.c_int), ("iters", C.c_int)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.orc_search.restype = C.c_int64
        _lib.orc_search.argtypes = [_vp, _vp, C.c_int, _vp, C.c_int, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_int, C.c_int, _vp, _vp]
        _lib.orc_brute.restype = C.c_int64
        _lib.orc_brute.argtypes = [_vp, _vp, C.c_int, _vp, C.c_int, _vp, _vp]
        _lib.orc_grid_params.argtypes = [C.c_float] * 5 + [_vp] * 4
        _lib.orc_pinv3x3.argtypes = [_vp, C.c_double, _vp]
        _lib.orc_pinv2x2.argtypes = [_vp, C.c_int, _vp, _vp]
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_get_threads.restype = C.c_int
        _lib.orc_norm_rows.argtypes = [_vp, C.c_int, C.c_double] + [_vp] * 9
        _lib.orc_dplus_shift.argtypes = [_vp, _vp, C.c_double] + [_vp] * 7
        _lib.orc_dplus_pressure.argtypes = [_vp] * 6 + [C.c_uint, _vp]
        for name in ("orc_W", "orc_dW", "orc_dWs"):
            fn = getattr(_lib, name)
            fn.restype = C.c_double
            fn.argtypes = [C.c_double, C.c_double]
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return int(lib().orc_get_threads())


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class CSR:
    """CSR pair list (int64 rowptr, int32 col), rows sorted ascending."""

    def __init__(self, rowptr, col):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.nrows = len(self.rowptr) - 1

    def c(self):
        return OrcCSR(self.nrows, _p(self.rowptr), _p(self.col))

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def pairs(self):
        rows = np.repeat(np.arange(self.nrows, dtype=np.int64), np.diff(self.rowptr))
        return rows, self.col.astype(np.int64)


def grid_params(vmin, vmax, hmax):
    qx, qy = C.c_float(), C.c_float()
    cx, cy = C.c_int(), C.c_int()
    lib().orc_grid_params(float(vmin[0]), float(vmin[1]), float(vmax[0]), float(vmax[1]),
                          float(hmax), C.byref(qx), C.byref(qy), C.byref(cx), C.byref(cy))
    return np.float32(qx.value), np.float32(qy.value), cx.value, cy.value


def search(xq, hq, xp, qmin, hmax, cells_x, radius=1, self_mode=False) -> CSR:
    """Reference pair list of queries xq (supports hq) against points xp."""
    xq, hq, xp = _f32(xq), _f32(hq), _f32(xp)
    nq, np_ = len(xq), len(xp)
    rowptr = np.zeros(nq + 1, dtype=np.int64)
    args = (_p(xq), _p(hq), nq, _p(xp), np_, float(qmin[0]), float(qmin[1]), float(hmax),
            int(cells_x), int(radius), 1 if self_mode else 0)
    nnz = lib().orc_search(*args, _p(rowptr), None)
    col = np.empty(max(nnz, 1), dtype=np.int32)
    lib().orc_search(*args, _p(rowptr), _p(col))
    return CSR(rowptr, col[:nnz])


def brute(xq, hq, xp) -> CSR:
    xq, hq, xp = _f32(xq), _f32(hq), _f32(xp)
    nq = len(xq)
    rowptr = np.zeros(nq + 1, dtype=np.int64)
    nnz = lib().orc_brute(_p(xq), _p(hq), nq, _p(xp), len(xp), _p(rowptr), None)
    col = np.empty(max(nnz, 1), dtype=np.int32)
    lib().orc_brute(_p(xq), _p(hq), nq, _p(xp), len(xp), _p(rowptr), _p(col))
    return CSR(rowptr, col[:nnz])


def transpose(a: CSR, ncols: int) -> CSR:
    rows, cols = a.pairs()
    order = np.lexsort((rows, cols))
    rp = np.zeros(ncols + 1, dtype=np.int64)
    np.add.at(rp, cols + 1, 1)
    return CSR(np.cumsum(rp), rows[order].astype(np.int32))


def phys_struct(p: dict) -> OrcPhys:
    s = OrcPhys()
    for k in ("h", "rho0", "c0", "gamma", "kappa", "alpha", "delta", "nu", "eps", "gx", "gy",
              "dt", "beta", "omega", "tau", "fluid_omega"):
        setattr(s, k, float(p.get(k, 0.0)))
    s.min_iters = int(p.get("min_iters", 2))
    s.max_iters = int(p.get("max_iters", 256))
    s.eos = int(p.get("eos", 0))
    return s


class _Keep:
    """Holds numpy arrays alive while their pointers are in a struct."""

    def __init__(self):
        self.refs = []

    def __call__(self, a):
        if a is not None:
            self.refs.append(a)
        return _p(a)


def fluid_struct(f: dict, keep: _Keep) -> OrcFluid:
    n = len(f["x"])
    return OrcFluid(n, keep(_f64(f["x"])), keep(_f64(f.get("v"))), keep(_f64(f["rho"])),
                    keep(_f64(f["area"])), keep(_f64(f.get("rest"))), keep(_f64(f.get("vol"))))


def bdy_struct(b: dict | None, keep: _Keep) -> OrcBdy | None:
    if b is None:
        return None
    n = len(b["x"])
    return OrcBdy(n, keep(_f64(b["x"])), keep(_f64(b.get("v"))), keep(_f64(b.get("rho"))),
                  keep(_f64(b.get("vol"))), keep(_f64(b.get("p"))), keep(_f64(b.get("rest"))),
                  keep(_f64(b.get("gx"))))


def _byref(s):
    return None if s is None else C.byref(s)


def fluid_density(f, phys, ff: CSR, b=None, fb: CSR | None = None):
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    n = F.n
    rho, vol = np.zeros(n), np.zeros(n)
    P = phys_struct(phys)
    lib().orc_fluid_density(C.byref(F), C.byref(P), C.byref(ff.c()), _byref(B),
                            _byref(fb.c() if fb is not None else None), _p(rho), _p(vol))
    return rho, vol


def dsph_rhs(f, phys, ff: CSR, b=None, fb: CSR | None = None, terms: int = 0):
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    n = F.n
    dpdt, acc, grho = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 2))
    P = phys_struct(phys)
    lib().orc_dsph_rhs(C.byref(F), C.byref(P), C.byref(ff.c()), _byref(B),
                       _byref(fb.c() if fb is not None else None), C.c_uint(terms), _p(dpdt),
                       _p(acc), _p(grho))
    return dpdt, acc, grho


def mdbc(f, phys, b, gf: CSR):
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    nb = B.n
    rho, p, cnt = np.zeros(nb), np.zeros(nb), np.zeros(nb, dtype=np.int32)
    P = phys_struct(phys)
    lib().orc_mdbc(C.byref(F), C.byref(P), C.byref(B), C.byref(gf.c()), _p(rho), _p(p), _p(cnt))
    return rho, p, cnt


def akinci_bdensity(f, phys, b, term, bf: CSR):
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    rho = np.zeros(B.n)
    term = keep(_f64(term))
    P = phys_struct(phys)
    lib().orc_akinci_bdensity(C.byref(F), C.byref(P), C.byref(B), term, C.byref(bf.c()), _p(rho))
    return rho


def pinv3x3(A, rtol=3 * np.finfo(np.float32).eps):
    A = _f64(A).reshape(9)
    X = np.zeros(9)
    lib().orc_pinv3x3(_p(A), float(rtol), _p(X))
    return X.reshape(3, 3)


def pinv2x2(M):
    M = _f64(M).reshape(-1, 4)
    n = len(M)
    L, e = np.zeros((n, 4)), np.zeros((n, 2))
    lib().orc_pinv2x2(_p(M), n, _p(L), _p(e))
    return L.reshape(n, 2, 2), e


def eos(phys, rho):
    rho = _f64(rho)
    p = np.zeros_like(rho)
    P = phys_struct(phys)
    lib().orc_eos(C.byref(P), _p(rho), len(rho), _p(p))
    return p


def divergence_phys(phys: dict) -> dict:
    """The limits divergenceSolve runs with (dfsph.py:240): divergence
    threshold and iteration bounds in place of the density solve's."""
    return dict(phys, tau=phys.get("div_tau", 1e-2), min_iters=phys.get("div_min_iters", 2),
                max_iters=phys.get("div_max_iters", 8))


def dfsph(f, phys, ff: CSR, b, fb: CSR, bf: CSR, acc, fixed_iters=0, max_hist=512, density=True):
    """densitySolve (density=True) or divergenceSolve (density=False, run
    with divergence_phys(phys)) of modules/dfsph.py:195-311."""
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    n, nb = F.n, (B.n if B is not None else 0)
    out = {k: np.zeros(n) for k in ("alpha", "src", "p2")}
    out.update(apred=np.zeros((n, 2)), vpred=np.zeros((n, 2)))
    out.update({k: np.zeros(max(nb, 1)) for k in ("b_alpha", "b_src", "b_p")})
    out["err_hist"] = np.zeros(max_hist)
    O = OrcDfsphOut(*[_p(out[k]) for k in ("alpha", "src", "p2", "apred", "vpred", "b_alpha",
                                            "b_src", "b_p", "err_hist")], max_hist, 0)
    acc = keep(_f64(acc))
    P = phys_struct(phys)
    lib().orc_dfsph_solve(C.byref(F), C.byref(P), C.byref(ff.c()), _byref(B),
                          _byref(fb.c() if fb is not None else None),
                          _byref(bf.c() if bf is not None else None), acc, int(fixed_iters),
                          1 if density else 0, C.byref(O))
    out["iters"] = O.iters
    out["err_hist"] = out["err_hist"][:min(O.iters, max_hist)]
    for k in ("b_alpha", "b_src", "b_p"):
        out[k] = out[k][:nb]
    return out



# ------------------------------------------------------ delta-plus shifting
def _csr_ref(c: CSR | None):
    return _byref(c.c() if c is not None else None)


def norm_rows(xq, h, a: CSR | None, xa, va, b: CSR | None = None, xb=None, vb=None):
    """M, L, eig of arbitrary query rows over two lists (densityDiffusion.py:
    30-43 for the boundary rows, solidBoundary.py:676-693)."""
    xq = _f64(xq)
    nq = len(xq)
    M, Lm, e = np.zeros((nq, 4)), np.zeros((nq, 4)), np.zeros((nq, 2))
    keep = _Keep()
    lib().orc_norm_rows(keep(xq), nq, float(h), _csr_ref(a), keep(_f64(xa)), keep(_f64(va)),
                        _csr_ref(b), keep(_f64(xb)), keep(_f64(vb)), _p(M), _p(Lm), _p(e))
    return M.reshape(nq, 2, 2), Lm.reshape(nq, 2, 2), e


def fluid_norm(f, phys, ff: CSR, b=None, fb: CSR | None = None):
    """M, L, eig of the fluid rows with the boundary part (deltaPlus.py:116-131)."""
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    n = F.n
    M, Lm, e = np.zeros((n, 4)), np.zeros((n, 4)), np.zeros((n, 2))
    P = phys_struct(phys)
    lib().orc_fluid_norm(C.byref(F), C.byref(P), C.byref(ff.c()), _byref(B), _csr_ref(fb), _p(M),
                         _p(Lm), _p(e))
    return M.reshape(n, 2, 2), Lm.reshape(n, 2, 2), e


def dplus_normal(f, phys, ff: CSR, Lm, eig, bx=None, beig=None, bvol_n=None, fb: CSR | None = None):
    """lambda gradient and fluid normal (deltaPlus.py:132-163)."""
    keep = _Keep()
    F = fluid_struct(f, keep)
    n = F.n
    grad, normal = np.zeros((n, 2)), np.zeros((n, 2))
    P = phys_struct(phys)
    lib().orc_dplus_normal(C.byref(F), C.byref(P), C.byref(ff.c()), keep(_f64(Lm)),
                           keep(_f64(eig)), keep(_f64(bx)), keep(_f64(beig)), keep(_f64(bvol_n)),
                           _csr_ref(fb), _p(grad), _p(normal))
    return grad, normal


def dplus_surface(f, phys, ff: CSR, normal, eig, bx=None, fb: CSR | None = None):
    """detectFluidSurface (deltaPlus.py:165-209): (cmax, mask, surface)."""
    keep = _Keep()
    F = fluid_struct(f, keep)
    n = F.n
    cmax, mask, surf = np.zeros(n), np.zeros(n), np.zeros(n)
    P = phys_struct(phys)
    lib().orc_dplus_surface(C.byref(F), C.byref(P), C.byref(ff.c()), keep(_f64(normal)),
                            keep(_f64(eig)), keep(_f64(bx)), _csr_ref(fb), _p(cmax), _p(mask),
                            _p(surf))
    return cmax, mask, surf


def dplus_shift(f, phys, packing, ff: CSR, normal, surface, b=None, fb: CSR | None = None):
    """computeShiftAmount + adjustShiftingAmount (deltaPlus.py:81-115,
    211-236): (raw shift, adjusted shift)."""
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    n = F.n
    raw, shift = np.zeros((n, 2)), np.zeros((n, 2))
    P = phys_struct(phys)
    lib().orc_dplus_shift(C.byref(F), C.byref(P), float(packing), C.byref(ff.c()), _byref(B),
                          _csr_ref(fb), keep(_f64(normal)), keep(_f64(surface)), _p(raw),
                          _p(shift))
    return raw, shift


def dplus_pressure(f, phys, ff: CSR, surface, b=None, fb: CSR | None = None,
                   terms: int = T_PRESS | T_PRESS_B):
    """computePressureAccelDeltaPlus (pressure.py:40-52, solidBoundary.py:882-887)."""
    keep = _Keep()
    F, B = fluid_struct(f, keep), bdy_struct(b, keep)
    acc = np.zeros((F.n, 2))
    P = phys_struct(phys)
    lib().orc_dplus_pressure(C.byref(F), C.byref(P), C.byref(ff.c()), _byref(B), _csr_ref(fb),
                             keep(_f64(surface)), int(terms), _p(acc))
    return acc

# ------------------------------------------------------------ full steps
DSPH_TERMS_MOG = (T_CONT | T_CONT_B | T_MOG | T_PRESS | T_PRESS_B | T_VISC | T_VISC_B | T_LAM |
                  T_GRAV)
DSPH_TERMS_DD = (DSPH_TERMS_MOG & ~T_MOG) | T_DDSPH


def dsph_terms(phys: dict) -> int:
    t = DSPH_TERMS_DD if phys.get("diffusion") == "deltaSPH" else DSPH_TERMS_MOG
    if not phys.get("boundary_diffusion", True):
        t &= ~T_VISC_B
    return t


def semi_implicit_step(x0, v0, acc, dt):
    """semiImplicitEuler (simulationBase.py:647-649): integrateValues with
    dudt = acc and dxdt = v0 + dt acc, i.e. v1 = v0 + dt a,
    x1 = x0 + dt (v0 + dt a)."""
    return x0 + dt * (v0 + dt * acc), v0 + dt * acc


class SceneOracle:
    """FP64 restatement of one scene's step (search in fp32 on the rounded
    positions, as the reference searches its fp32 state)."""

    def __init__(self, scene):
        self.sc = scene
        self.qx, self.qy, self.cx, self.cy = grid_params(scene.vmin, scene.vmax, scene.support)
        self.hmax = np.float32(scene.support)
        n, nb = scene.n, scene.nb
        self.h = np.full(n, scene.support, np.float32)
        self.bh = np.full(nb, scene.support, np.float32)
        self.area = np.full(n, scene.area)
        self.rest = np.full(n, scene.rho0)
        self.brest = np.full(nb, scene.rho0)

    def pairs(self, x):
        sc = self.sc
        xf = np.asarray(x, dtype=np.float32)
        q = (self.qx, self.qy)
        ff = search(xf, self.h, xf, q, self.hmax, self.cx, 1, self_mode=True)
        bf = search(sc.bx, self.bh, xf, q, self.hmax, self.cx, 1)
        gf = search(sc.bgx, self.bh, xf, q, self.hmax, self.cx, 1) if sc.bgx is not None else None
        return ff, bf, gf, transpose(bf, len(xf))

    def dsph_rhs(self, x, v, rho):
        sc, P = self.sc, self.sc.phys
        ff, bf, gf, fb = self.pairs(x)
        f = dict(x=x, v=v, rho=rho, area=self.area, rest=self.rest)
        brho, bp, _ = mdbc(f, P, dict(x=sc.bx, gx=sc.bgx, vol=sc.bvol), gf)
        b = dict(x=sc.bx, rho=brho, vol=sc.bvol, p=bp)
        dpdt, acc, _ = dsph_rhs(f, P, ff, b, fb, dsph_terms(P))
        return dpdt, acc

    def dsph_step(self, x, v, rho):
        """RK4 step (simulationBase.py:679-705)."""
        dt = self.sc.phys["dt"]
        r0 = self.sc.rho0
        x0, v0, p0 = (np.asarray(a, dtype=np.float64) for a in (x, v, rho))
        ks = []
        cx, cv, cr = x0, v0, p0
        for c in (0.5, 0.5, 1.0, None):
            dpdt, acc = self.dsph_rhs(cx, cv, cr)
            ks.append((acc, cv, dpdt))
            if c is not None:
                cv, cx, cr = v0 + c * dt * acc, x0 + c * dt * cv, p0 + c * dt * dpdt / r0
        w = (1, 2, 2, 1)
        ka = sum(wi * k[0] for wi, k in zip(w, ks))
        kv = sum(wi * k[1] for wi, k in zip(w, ks))
        kr = sum(wi * k[2] for wi, k in zip(w, ks))
        return x0 + dt / 6 * kv, v0 + dt / 6 * ka, p0 + dt / 6 * kr / r0

    # ------------------------------------------------------- delta-plus
    def boundary_pairs(self):
        """Static boundary->boundary pairs (solidBoundary.py:180), brute force."""
        if getattr(self, "_bb", None) is None:
            bx = np.asarray(self.sc.bx, dtype=np.float32)
            self._bb = brute(bx, self.bh, bx)
        return self._bb

    def dplus_surface(self, x, v, rho, lists=None):
        """detectSurface of one right-hand side (deltaSPH.py:146-169 then
        deltaPlus.py:242-245): returns (f, b, lists, eig, beig, normal,
        surface)."""
        sc, P = self.sc, self.sc.phys
        ff, bf, gf, fb = lists if lists is not None else self.pairs(x)
        f = dict(x=x, v=v, rho=rho, area=self.area, rest=self.rest)
        brho, bp, _ = mdbc(f, P, dict(x=sc.bx, gx=sc.bgx, vol=sc.bvol), gf)
        b = dict(x=sc.bx, rho=brho, vol=sc.bvol, p=bp)
        _, Lm, eig = fluid_norm(f, P, ff, b, fb)
        vf = self.area / np.asarray(rho, dtype=np.float64)
        _, _, beig = norm_rows(sc.bx, P["h"], bf, x, vf, self.boundary_pairs(), sc.bx,
                               sc.bvol / brho)
        n, nb = len(x), len(sc.bx)
        m = min(n, nb)
        bvol_n = np.zeros(nb)
        bvol_n[:m] = vf[:m]          # deltaPlus.py:151 indexes the fluid arrays by b
        _, normal = dplus_normal(f, P, ff, Lm, eig, sc.bx, beig, bvol_n, fb)
        _, _, surf = dplus_surface(f, P, ff, normal, eig, sc.bx, fb)
        return f, b, (ff, bf, gf, fb), eig, beig, normal, surf

    def dplus_rhs(self, x, v, rho):
        """delta-plus right-hand side: the delta-SPH terms with the pressure
        switch of pressure.py:40-52.  Returns dpdt, acc, normal, surface."""
        P = self.sc.phys
        f, b, (ff, bf, gf, fb), _, _, normal, surf = self.dplus_surface(x, v, rho)
        terms = dsph_terms(P) & ~(T_PRESS | T_PRESS_B)
        dpdt, acc, _ = dsph_rhs(f, P, ff, b, fb, terms)
        acc = acc + dplus_pressure(f, P, ff, surf, b, fb)
        return dpdt, acc, normal, surf

    def dplus_step(self, x, v, rho, packing, shifting=True):
        """RK4 step of the delta-plus scheme, then the shift
        (simulationBase.py:679-732) with the last stage's normal and mask."""
        sc, P = self.sc, self.sc.phys
        dt, r0 = P["dt"], sc.rho0
        x0, v0, p0 = (np.asarray(a, dtype=np.float64) for a in (x, v, rho))
        ks = []
        cx, cv, cr = x0, v0, p0
        for c in (0.5, 0.5, 1.0, None):
            dpdt, acc, normal, surf = self.dplus_rhs(cx, cv, cr)
            ks.append((acc, cv, dpdt))
            if c is not None:
                cv, cx, cr = v0 + c * dt * acc, x0 + c * dt * cv, p0 + c * dt * dpdt / r0
        w = (1, 2, 2, 1)
        ka = sum(wi * k[0] for wi, k in zip(w, ks))
        kv = sum(wi * k[1] for wi, k in zip(w, ks))
        kr = sum(wi * k[2] for wi, k in zip(w, ks))
        x1, v1, r1 = x0 + dt / 6 * kv, v0 + dt / 6 * ka, p0 + dt / 6 * kr / r0
        if shifting:
            ff, bf, gf, fb = self.pairs(x1)
            f = dict(x=x1, v=v1, rho=r1, area=self.area, rest=self.rest)
            brho, bp, _ = mdbc(f, P, dict(x=sc.bx, gx=sc.bgx, vol=sc.bvol), gf)
            b = dict(x=sc.bx, rho=brho, vol=sc.bvol, p=bp)
            _, sh = dplus_shift(f, P, packing, ff, normal, surf, b, fb)
            x1 = x1 + sh
        return x1, v1, r1

    def dfsph_step(self, x, v, fixed_iters=0):
        """Semi-implicit Euler DFSPH step (dfsph.py:134-202,
        simulationBase.py:647-652).  Returns x, v, rho, solver output."""
        sc, P = self.sc, self.sc.phys
        x0, v0 = np.asarray(x, dtype=np.float64), np.asarray(v, dtype=np.float64)
        ff, bf, gf, fb = self.pairs(x0)
        f = dict(x=x0, v=v0, rho=np.ones(len(x0)), area=self.area, rest=self.rest)
        b = dict(x=sc.bx, vol=sc.bvol, rest=self.brest)
        rho, vol = fluid_density(f, P, ff, b, fb)
        brho = akinci_bdensity(f, P, b, sc.bterm, bf)
        f["rho"] = rho
        b["rho"] = brho
        acc = np.tile([P["gx"], P["gy"]], (len(x0), 1)).astype(np.float64)
        div = None
        if P.get("divergence", False):
            # dfsph.py:175-179: the divergence solve first, its pressure
            # acceleration added to the acceleration the density solve
            # predicts velocities with
            div = dfsph(f, divergence_phys(P), ff, b, fb, bf, acc, density=False)
            acc = acc + div["apred"]
        out = dfsph(f, P, ff, b, fb, bf, acc, fixed_iters=fixed_iters)
        out["divergence"] = div
        acc = acc + out["apred"]
        fv = dict(f, vol=vol)
        _, visc, _ = dsph_rhs(fv, P, ff, None, None, T_VISC | T_LAM)
        acc = acc + visc
        x1, v1 = semi_implicit_step(x0, v0, acc, P["dt"])
        out["rho"], out["vol"], out["brho"], out["acc"] = rho, vol, brho, acc
        return x1, v1, rho, out
