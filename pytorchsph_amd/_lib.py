"""ctypes binding of libsph_mi355x.so, the gfx950 C-ABI library.

The structures below mirror include/sph_mi355x.h field by field.  Every call
goes through `call()`, which raises RuntimeError with the library's message
on a non-zero status.  The library must exist: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np
import torch

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsph_mi355x.so"

c_vp = C.c_void_p
c_int = C.c_int


class SphRk4Stage(C.Structure):
    _fields_ = [("stage", c_int), ("x0", c_vp), ("v0", c_vp), ("r0", c_vp), ("area", c_vp),
                ("K", c_vp), ("x_out", c_vp), ("v_out", c_vp), ("rho_out", c_vp),
                ("acc_out", c_vp), ("frec_next", c_vp)]


PERMUTE_MAX_FIELDS = 8


class SphFields(C.Structure):
    _fields_ = [("ptr", c_vp * PERMUTE_MAX_FIELDS), ("width", c_int * PERMUTE_MAX_FIELDS),
                ("nfields", c_int)]


class SphGrid(C.Structure):
    _fields_ = [
        ("scenes", c_vp), ("pscene", c_vp), ("x", c_vp),
        ("np", c_int), ("nscenes", c_int), ("ncells", c_int),
        ("cell_count", c_vp), ("cell_start", c_vp), ("key", c_vp), ("slot", c_vp),
        ("sorted_idx", c_vp), ("sorted_x", c_vp), ("status", c_vp),
    ]


class SphQuery(C.Structure):
    _fields_ = [
        ("x", c_vp), ("h", c_vp), ("qscene", c_vp),
        ("nq", c_int), ("radius", c_int), ("self_sorted", c_int), ("qcell", c_vp),
    ]


class SphFluid(C.Structure):
    _fields_ = [
        ("x", c_vp), ("v", c_vp), ("rho", c_vp), ("area", c_vp), ("rest", c_vp),
        ("vol", c_vp), ("pscene", c_vp), ("np", c_int),
    ]


class SphBoundary(C.Structure):
    _fields_ = [
        ("x", c_vp), ("v", c_vp), ("rho", c_vp), ("vol", c_vp), ("p", c_vp), ("rest", c_vp),
        ("bscene", c_vp), ("nb", c_int),
    ]


class SphCSR(C.Structure):
    _fields_ = [("rowptr", c_vp), ("col", c_vp), ("nrows", c_int)]


class SphStencilList(C.Structure):
    _fields_ = [("runs", c_vp), ("masks", c_vp), ("n", c_int)]


DFSPH_FIELDS = ("vpred", "aact", "alpha", "src", "p", "p2", "apred", "resid", "xrec", "arec",
                "prec", "wvec", "b_aact", "b_alpha", "b_src", "b_p", "b_resid", "bxrec", "err_acc", "err_hist",
                "iters", "active", "any_active", "status", "nfluid", "nbdy", "fbegin", "bbegin",
                "act_scene", "act_pre", "act_n")


class SphDfsph(C.Structure):
    _fields_ = [(n, c_vp) for n in DFSPH_FIELDS] + [
        ("nscenes", c_int), ("max_hist", c_int), ("grid_blocks", c_int), ("divergence", c_int)]


# device-array record types (numpy layouts identical to the C structs)
GRID_DESC_DTYPE = np.dtype([
    ("qmin_x", "<f4"), ("qmin_y", "<f4"), ("hmax", "<f4"),
    ("cells_x", "<i4"), ("cells_y", "<i4"), ("cell_begin", "<i4"),
    ("p_begin", "<i4"), ("p_count", "<i4"),
])
PHYS_DTYPE = np.dtype([
    ("h", "<f4"), ("rho0", "<f4"), ("c0", "<f4"), ("gamma", "<f4"), ("kappa", "<f4"),
    ("alpha", "<f4"), ("delta", "<f4"), ("nu", "<f4"), ("eps", "<f4"), ("gx", "<f4"),
    ("gy", "<f4"), ("dt", "<f4"), ("beta", "<f4"), ("omega", "<f4"), ("tau", "<f4"),
    ("fluid_omega", "<f4"), ("min_iters", "<i4"), ("max_iters", "<i4"), ("eos", "<i4"),
    ("pad", "<i4"),
])
assert GRID_DESC_DTYPE.itemsize == 32 and PHYS_DTYPE.itemsize == 80

# status bits and gather terms (sph_mi355x.h)
ST_OUTSIDE, ST_OVERFLOW, ST_NAN, ST_ITERCAP = 1, 2, 4, 8
T_CONT, T_CONT_B, T_MOG, T_DDSPH = 1 << 0, 1 << 1, 1 << 2, 1 << 3
T_PRESS, T_PRESS_B, T_VISC, T_VISC_B = 1 << 4, 1 << 5, 1 << 6, 1 << 7
T_LAM, T_GRAV, T_ACCUM = 1 << 8, 1 << 9, 1 << 16
TUNE_SEARCH_LANES, TUNE_GATHER_LANES, TUNE_SOLVE_LANES, TUNE_FILL_LANES, TUNE_WALK_LANES = \
    0, 1, 2, 3, 4
TUNE_BLOCK = 5
DFSPH_ERR_SLOTS = 64


def dfsph_err_stride(nscenes: int) -> int:
    """Row stride (doubles) of SphDfsph.err_acc: SPH_DFSPH_ERR_STRIDE."""
    return (nscenes + 15) // 16 * 16
TUNE_KNOBS = {"search": TUNE_SEARCH_LANES, "gather": TUNE_GATHER_LANES, "solve": TUNE_SOLVE_LANES,
              "fill": TUNE_FILL_LANES, "walk": TUNE_WALK_LANES, "block": TUNE_BLOCK}

# exported entry points: name -> argtypes (restype is int unless noted)
_P = C.POINTER
_SIGS = {
    "sph_last_error": ([], C.c_char_p),
    "sph_version": ([], c_int),
    "sph_scan_workspace_bytes": ([c_int], C.c_size_t),
    "sph_grid_workspace_bytes": ([c_int], C.c_size_t),
    "sph_scan_exclusive": ([c_vp, c_int, c_vp, c_vp, C.c_size_t, c_vp], c_int),
    "sph_scan_exclusive_noalloc": ([c_vp, c_int, c_vp, c_vp], c_int),
    "sph_grid_build_keyed": ([_P(SphGrid), c_vp, c_vp, C.c_size_t, c_vp], c_int),
    "sph_grid_sort_workspace_bytes": ([c_int, c_int], C.c_size_t),
    "sph_grid_build_sorted": ([_P(SphGrid), c_vp, c_vp, C.c_size_t, c_vp], c_int),
    "sph_pair_thresholds": ([c_vp, c_int, c_vp, c_vp], c_int),
    "sph_fill_i32": ([c_vp, c_int, c_int, c_vp], c_int),
    "sph_fill_f32": ([c_vp, c_int, C.c_float, c_vp], c_int),
    "sph_set_tuning": ([c_int, c_int], c_int),
    "sph_get_tuning": ([c_int], c_int),
    "sph_grid_build": ([_P(SphGrid), c_vp, C.c_size_t, c_vp], c_int),
    "sph_neighbor_count": ([_P(SphGrid), _P(SphQuery), c_vp, c_vp], c_int),
    "sph_neighbor_fill": ([_P(SphGrid), _P(SphQuery), c_vp, c_vp, C.c_int64, c_vp], c_int),
    "sph_csr_sort_rows": ([_P(SphCSR), c_vp, c_vp], c_int),
    "sph_stencil_list_build": ([_P(SphGrid), c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_iota": ([c_vp, c_int, c_vp], c_int),
    "sph_fluid_density_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphBoundary),
                              _P(SphCSR), c_vp, c_vp, c_vp], c_int),
    "sph_dfsph_alpha_source_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphBoundary),
                                   _P(SphCSR), _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_accel_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphCSR), _P(SphDfsph),
                            c_vp], c_int),
    "sph_dfsph_pressure_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphBoundary),
                               _P(SphCSR), _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_iterate_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphBoundary),
                              _P(SphCSR), _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_finish_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphCSR), _P(SphDfsph),
                             c_vp, c_vp], c_int),
    "sph_dfsph_finish_visc_sl": ([_P(SphFluid), c_vp, _P(SphStencilList), _P(SphCSR),
                                  _P(SphDfsph), c_vp, C.c_uint, c_vp, c_vp], c_int),
    "sph_csr_transpose": ([_P(SphCSR), c_int, c_vp, c_vp, c_vp, c_vp, c_vp, C.c_size_t, c_vp], c_int),
    "sph_pair_geometry": ([_P(SphCSR), c_vp, c_vp, c_vp, C.c_float, C.c_float, c_vp, c_vp, c_vp,
                           c_vp, c_vp], c_int),
    "sph_fluid_density": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR), c_vp, c_vp,
                           c_vp, c_vp], c_int),
    "sph_dsph_gather": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR), c_vp,
                         C.c_uint, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dsph_normalization": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR), c_vp,
                                c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dsph_density_gradient": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR),
                                   c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_pinv2x2": ([c_vp, c_int, c_vp, c_vp, c_vp], c_int),
    "sph_mdbc_density": ([_P(SphFluid), c_vp, c_vp, c_vp, c_vp, c_int, _P(SphCSR), c_vp, c_vp,
                          c_vp, c_vp], c_int),
    "sph_mdbc_density_walk": ([_P(SphFluid), c_vp, c_vp, c_vp, c_vp, c_vp, c_int, _P(SphGrid),
                               c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_mdbc_density_walk_records": ([_P(SphFluid), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                       c_int, _P(SphGrid), c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_eos": ([c_vp, c_vp, c_vp, c_int, c_vp, c_vp], c_int),
    "sph_akinci_boundary_density": ([_P(SphFluid), c_vp, c_vp, c_vp, c_int, c_vp, _P(SphCSR),
                                     c_vp, c_vp], c_int),
    "sph_dfsph_prepare": ([_P(SphFluid), c_vp, c_vp, _P(SphBoundary), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_alpha_source": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR),
                                _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_iterate": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR),
                           _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_accel": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_pressure": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR),
                            _P(SphCSR), _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_control": ([c_vp, _P(SphDfsph), c_vp], c_int),
    "sph_dfsph_finish": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphCSR), _P(SphDfsph), c_vp, c_vp],
                         c_int),
    "sph_dfsph_finish_visc": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphCSR), _P(SphDfsph), c_vp,
                               C.c_uint, c_vp, c_vp], c_int),
    "sph_dsph_pack": ([_P(SphFluid), c_vp, _P(SphBoundary), c_vp, c_vp, c_vp], c_int),
    "sph_dsph_gather_packed": ([c_vp, c_vp, c_vp, c_int, c_vp, _P(SphCSR), _P(SphCSR), c_vp,
                                C.c_uint, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dsph_gather_packed_dplus": ([c_vp, c_vp, c_vp, c_int, c_vp, _P(SphCSR), _P(SphCSR), c_vp,
                                      C.c_uint, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dsph_gather_rk4": ([c_vp, c_vp, c_vp, c_int, c_vp, _P(SphCSR), _P(SphCSR), c_vp,
                             C.c_uint, c_vp, _P(SphRk4Stage), c_vp], c_int),
    "sph_dsph_cellwalk": ([_P(SphGrid), c_vp, c_int, c_vp, c_vp, c_vp, _P(SphCSR), _P(SphGrid),
                           C.c_uint, c_vp, c_vp, c_vp], c_int),
    "sph_dsph_cellwalk_rk4": ([_P(SphGrid), c_vp, c_int, c_vp, c_vp, c_vp, _P(SphCSR),
                               _P(SphGrid), C.c_uint, _P(SphRk4Stage), c_vp], c_int),
    "sph_integrate_rk4": ([c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_integrate_semi_implicit": ([c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                     c_vp, c_vp, c_vp], c_int),
    "sph_permute_f32": ([c_vp, c_vp, c_int, c_int, c_vp, c_vp], c_int),
    "sph_permute_fields": ([_P(SphFields), c_vp, c_int, c_vp, c_vp], c_int),
    "sph_normalization_rows": ([c_vp, c_vp, c_int, c_vp, _P(SphCSR), c_vp, c_vp, _P(SphCSR), c_vp,
                                c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dplus_normal": ([_P(SphFluid), c_vp, _P(SphCSR), c_vp, c_vp, c_vp, c_vp, c_vp,
                          _P(SphCSR), c_vp, c_vp, c_vp], c_int),
    "sph_dplus_surface": ([_P(SphFluid), c_vp, _P(SphCSR), c_vp, c_vp, c_vp, _P(SphCSR), c_vp,
                           c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dplus_detect": ([_P(SphFluid), c_vp, _P(SphCSR), c_vp, c_vp, c_vp, c_vp, c_vp,
                          _P(SphCSR), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dplus_shift": ([_P(SphFluid), c_vp, c_int, c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR),
                         c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "sph_dplus_pressure": ([_P(SphFluid), c_vp, _P(SphCSR), _P(SphBoundary), _P(SphCSR), c_vp,
                            C.c_uint, c_vp, c_vp], c_int),
}

_lib = None


def load() -> C.CDLL:
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m pytorchsph_amd.build` "
                "(there is no CPU fallback for the HIP path)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def call(name: str, *args) -> None:
    """Calls an entry point and raises on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.sph_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed (status {rc}): {msg}")


def set_tuning(**lanes: int) -> dict:
    """Sets launch-shape knobs by name: lanes per row (search=, gather=, solve=,
    fill=, walk=) and the workgroup size of the list gathers and per-query
    search (block= 64, 128 or 256); returns
    the previous values.  Host-side only: no device is touched."""
    old = get_tuning()
    for name, v in lanes.items():
        call("sph_set_tuning", TUNE_KNOBS[name], int(v))
    return old


def get_tuning() -> dict:
    lib = load()
    return {name: int(lib.sph_get_tuning(k)) for name, k in TUNE_KNOBS.items()}


DFSPH_GRID_CAP = 8192


def dfsph_grid_blocks(rows: int, cap: int | None = None) -> int:
    """Workgroups of the DFSPH iteration kernels: one lane group per row
    (fluid + boundary), capped at DFSPH_GRID_CAP; the kernels stride over
    larger work."""
    per_block = 256 // get_tuning()["solve"]   # rows per workgroup pass
    return max(1, min(-(-rows // per_block), cap or DFSPH_GRID_CAP))


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def csr(rowptr, col, nrows: int) -> SphCSR:
    return SphCSR(ptr(rowptr), ptr(col), int(nrows))


def records_to_tensor(arr: np.ndarray, device) -> torch.Tensor:
    """Uploads a numpy structured array as raw bytes."""
    raw = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).to(device)
