"""C-ABI boundary checks that need no GPU: the library loads and exports
every entry point include/sph_mi355x.h declares; record layouts match."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

from pytorchsph_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_declarations():
    assert declared_symbols() == set(L.exported_symbols())


def test_version_and_error_string():
    lib = L.load()
    assert lib.sph_version() >= 1
    assert isinstance(lib.sph_last_error(), bytes)


def test_argument_errors_are_reported_without_gpu():
    # a null grid is rejected on the host before any launch
    lib = L.load()
    rc = lib.sph_grid_build(None, None, 0, None)
    assert rc == 1
    assert b"null grid" in lib.sph_last_error()


def test_record_layouts():
    assert L.GRID_DESC_DTYPE.itemsize == 32
    assert L.PHYS_DTYPE.itemsize == 80
    assert C.sizeof(L.SphGrid) == 8 * 3 + 4 * 3 + 4 + 8 * 7
    # x, h, qscene; nq, radius, self_sorted; padding; qcell
    assert C.sizeof(L.SphQuery) == 8 * 3 + 4 * 3 + 4 + 8
    assert L.SphQuery.qcell.offset == 40
    assert C.sizeof(L.SphFields) == 8 * 8 + 4 * 8 + 4 + 4
    assert C.sizeof(L.SphRk4Stage) == 8 + 8 * 10


def test_scan_workspace_sizes():
    lib = L.load()
    assert lib.sph_scan_workspace_bytes(1) >= 8
    assert lib.sph_scan_workspace_bytes(10_000_000) >= 4 * (10_000_000 // 2048)


def test_bad_arguments_rejected_on_host():
    """Entry points validate their arguments before any launch: a null or
    inconsistent argument returns SPH_E_ARG with the function named in the
    message (no GPU is touched)."""
    import ctypes as C

    lib = L.load()
    cases = {
        "sph_neighbor_count": lambda: lib.sph_neighbor_count(None, None, None, None),
        "sph_dsph_gather_packed": lambda: lib.sph_dsph_gather_packed(
            None, None, None, 0, None, None, None, None, 0, None, None, None, None),
        "sph_dsph_gather_packed_dplus": lambda: lib.sph_dsph_gather_packed_dplus(
            None, None, None, 0, None, None, None, None, 0, None, None, None, None, None),
        "sph_normalization_rows": lambda: lib.sph_normalization_rows(
            None, None, 4, None, None, None, None, None, None, None, None, None, None, None),
        "sph_dplus_normal": lambda: lib.sph_dplus_normal(None, None, None, None, None, None, None,
                                                         None, None, None, None, None),
        "sph_dplus_surface": lambda: lib.sph_dplus_surface(None, None, None, None, None, None,
                                                           None, None, None, None, None, None),
        "sph_dplus_detect": lambda: lib.sph_dplus_detect(None, None, None, None, None, None, None,
                                                         None, None, None, None, None, None,
                                                         None, None, None),
        "sph_dplus_shift": lambda: lib.sph_dplus_shift(None, None, 1, None, None, None, None,
                                                       None, None, None, None, None),
        "sph_dplus_pressure": lambda: lib.sph_dplus_pressure(None, None, None, None, None, None,
                                                             0, None, None),
        "sph_permute_fields": lambda: lib.sph_permute_fields(None, None, 0, None, None),
        "sph_dsph_cellwalk": lambda: lib.sph_dsph_cellwalk(None, None, 0, None, None, None, None,
                                                           None, 0, None, None, None),
        "sph_dsph_cellwalk_rk4": lambda: lib.sph_dsph_cellwalk_rk4(None, None, 0, None, None,
                                                                   None, None, None, 0, None,
                                                                   None),
        "sph_mdbc_density_walk": lambda: lib.sph_mdbc_density_walk(None, None, None, None, None,
                                                                   None, 0, None, None, None,
                                                                   None, None),
        "sph_mdbc_density_walk_records": lambda: lib.sph_mdbc_density_walk_records(
            None, None, None, None, None, None, None, None, 0, None, None, None, None, None, None),
    }
    for name, call in cases.items():
        assert call() == 1, name
        assert name.encode() in lib.sph_last_error(), (name, lib.sph_last_error())
