"""torch.ops.sph: the reference's neighbour-search entry points as torch
operators over the C-ABI library (csrc/torch_ops.cpp).

    buildNeighborListCUDA(queryParticles, support, sortedParticles,
        sortedSupport, hashTable, cellIndices, cumCell, cellSpan, sort, qMin,
        hMax, cellsX, hashMapLength, searchRadius) -> [i, j]
    buildNeighborList(sortedParticles, sortedSupport, hashTable, cellIndices,
        cumCell, cellSpan, sort, cellsX, hashMapLength, searchRadius) -> [rows, cols]
    buildNeighborListUnsortedPerParticle(inputParticles, inputSupport,
        sortedParticles, sortedSupport, hashTable, cellIndices, cumCell,
        cellSpan, sort, cellsX, hashMapLength, qMin, hMax, searchRadius) -> [rows, cols]

with the argument lists of cppSrc/neighSearch.cpp:143-177,261-267.  Being
registered operators (TORCH_LIBRARY), they are callable from TorchScript,
which the reference's pybind module is not.  `neighborSearch` is the
namespace object that replaces the reference's JIT-built module at
src/modules/neighborSearch.py:28.  Importing this module loads the library
and raises if it has not been built: there is no fallback.
"""
from __future__ import annotations

from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libsph_torch.so"
_loaded = False


def load():
    global _loaded
    if not _loaded:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run pytorchsph_amd.build.build() first")
        torch.ops.load_library(str(LIB_PATH))
        _loaded = True
    return torch.ops.sph


neighborSearch = load()
