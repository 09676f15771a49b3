"""Builds libsph_mi355x.so, the gfx950 C-ABI library, from csrc/*.hip, and
libsph_torch.so, the torch operators over it, from csrc/torch_ops.cpp.

Each translation unit of the library is compiled to an object in parallel,
then linked into one shared library next to this file.  The library is
plain C-ABI (no torch headers), so it builds in seconds; the operator
library includes the torch headers and takes longer.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = REPO / "include"
OBJDIR = PKG / "_build"
LIBNAME = "libsph_mi355x.so"
LIB = PKG / LIBNAME

ARCH = os.environ.get("SPH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed fp32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no implicit fused multiply-add formation: a sum then has the same
    # bits in every kernel that evaluates it (multiply-adds are written out
    # with fmaf where the code wants them)
    "-ffp-contract=off",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = OBJDIR / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
    return obj


TORCH_LIBNAME = "libsph_torch.so"
TORCH_LIB = PKG / TORCH_LIBNAME


def build_torch_ops(verbose: bool = False) -> Path:
    """Compiles csrc/torch_ops.cpp, the torch operators (TORCH_LIBRARY(sph))
    over the C-ABI library, and links it against libsph_mi355x.so (found
    next to it through its rpath) and the installed torch.  Host code only:
    its device work is the library's."""
    import torch
    from torch.utils import cpp_extension as ce

    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    tlib = ce.library_paths()[0]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tmp = TORCH_LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
           *[f"-I{p}" for p in ce.include_paths()], f"-I{rocm}/include", f"-I{INCLUDE}",
           str(CSRC / "torch_ops.cpp"), "-o", str(tmp),
           f"-L{PKG}", "-lsph_mi355x", f"-L{tlib}", "-lc10", "-lc10_hip", "-ltorch_cpu",
           "-ltorch_hip", "-ltorch", f"-L{rocm}/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tlib}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"torch_ops.cpp failed:\n{res.stderr}")
    os.replace(tmp, TORCH_LIB)
    if verbose:
        print(f"built {TORCH_LIB}")
    return TORCH_LIB


def build(verbose: bool = False, torch_ops: bool = True) -> Path:
    """Compiles every csrc/*.hip for the target arch and links the library,
    then (torch_ops) the torch operator library over it."""
    OBJDIR.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    if torch_ops:
        build_torch_ops(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
