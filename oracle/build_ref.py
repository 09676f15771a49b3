"""Builds the reference's own CPU neighbour search into oracle/_ref/.

TEST INFRASTRUCTURE ONLY.  Compiles cppSrc/neighSearch.cpp of
wi-re/pytorchSPH from its source file where it lies under /root/reference
(nothing is copied into this repository) with g++ against the torch headers
of this image, into oracle/_ref/neighsearch_ref.so (git-ignored).  Only the
two CPU entry points are used: buildNeighborList (neighSearch.cpp:172-259)
and buildNeighborListUnsortedPerParticle (:261-331).  The file also declares
two CUDA implementations defined in neighSearch_cuda.cu, which cannot be
compiled here; they stay unresolved (lazy binding, never called).

Flags: -O2 -ffp-contract=off, i.e. cppSrc/setup.py's baseline-x86-64
arithmetic (no FMA contraction); the JIT build of neighborSearch.py:28 adds
-march=native, whose FMA contraction can move near-tie pairs.

Usage: python oracle/build_ref.py   (skips the build when up to date)
"""
from __future__ import annotations

import subprocess
import sys
import sysconfig
from pathlib import Path

REF_SRC = Path("/root/reference/cppSrc/neighSearch.cpp")
OUT_DIR = Path(__file__).resolve().parent / "_ref"
OUT = OUT_DIR / "neighsearch_ref.so"


def build(force: bool = False) -> Path | None:
    if not REF_SRC.exists():
        print("build_ref: reference sources absent; nothing to build")
        return None
    if OUT.exists() and not force and OUT.stat().st_mtime > REF_SRC.stat().st_mtime:
        return OUT
    import torch
    from torch.utils.cpp_extension import include_paths

    OUT_DIR.mkdir(exist_ok=True)
    torch_lib = Path(torch.__file__).parent / "lib"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-w", "-DTORCH_EXTENSION_NAME=neighsearch_ref", "-DTORCH_API_INCLUDE_EXTENSION_H",
           f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}"]
    cmd += [f"-I{p}" for p in include_paths()]
    cmd += [f"-I{sysconfig.get_paths()['include']}", str(REF_SRC), "-o", str(OUT),
            f"-L{torch_lib}", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python",
            "-Wl,-z,lazy", f"-Wl,-rpath,{torch_lib}"]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    out = build(force="--force" in sys.argv)
    if out:
        print(f"built {out}")
