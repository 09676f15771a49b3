"""Drives the reference-built CPU neighbour search (oracle/_ref, see
build_ref.py).  TEST INFRASTRUCTURE ONLY.

The reference prepares the sorted points and the cell hash table in torch
(sortPositions and constructHashMap, neighborSearch.py:30-88) before calling
its C++ search; `hash_structures` restates those steps on CPU tensors so
that the compiled functions receive exactly the inputs the reference hands
them.  The hash-table length only changes the table's layout, never the
pair set; the reference takes nextprime(N) (neighborSearch.py:95).
"""
from __future__ import annotations

import importlib.util
import os
import sys
from pathlib import Path

import numpy as np

SO = Path(__file__).resolve().parent / "_ref" / "neighsearch_ref.so"
_mod = None


def available() -> bool:
    return SO.exists()


def load():
    """Imports the compiled reference module with lazy symbol binding (its
    two CUDA entry points are never called and stay unresolved)."""
    global _mod
    if _mod is None:
        import torch  # noqa: F401  (libtorch symbols)
        flags = sys.getdlopenflags()
        sys.setdlopenflags(os.RTLD_LAZY)
        try:
            spec = importlib.util.spec_from_file_location("neighsearch_ref", SO)
            mod = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(mod)
        finally:
            sys.setdlopenflags(flags)
        _mod = mod
    return _mod


def _next_prime(n: int) -> int:
    def is_prime(k):
        if k < 2:
            return False
        i = 2
        while i * i <= k:
            if k % i == 0:
                return False
            i += 1
        return True
    k = n + 1
    while not is_prime(k):
        k += 1
    return k


def hash_structures(x, h, vmin, vmax):
    """Sorted points, cell table and hash table as sortPositions and
    constructHashMap build them (fp32 tensor arithmetic, supportScale 1)."""
    import torch

    xt = torch.from_numpy(np.ascontiguousarray(x, np.float32))
    ht = torch.from_numpy(np.ascontiguousarray(h, np.float32))
    hmax = torch.max(ht)
    qmin = torch.tensor(vmin, dtype=torch.float32) - hmax
    qmax = torch.tensor(vmax, dtype=torch.float32) + 2 * hmax
    cells = torch.ceil((qmax - qmin) / hmax).to(torch.int32)
    idx = torch.ceil((xt - qmin) / hmax).to(torch.int32)
    lin = idx[:, 0] + cells[0] * idx[:, 1]
    order = torch.argsort(lin)
    slin = lin[order]
    sx, sh = xt[order, :].contiguous(), ht[order].contiguous()
    cell_keys, counts = torch.unique_consecutive(slin, return_counts=True)
    counts = counts.to(torch.int32)
    starts = (torch.cumsum(counts, 0) - counts).to(torch.int32)
    P = _next_prime(len(x))
    cx = int(cells[0])
    hashed = ((cell_keys % cx) * 3 + torch.div(cell_keys, cx, rounding_mode="trunc") * 5) % P
    horder = torch.argsort(hashed)
    hkeys, hcounts = torch.unique_consecutive(hashed[horder], return_counts=True)
    hcounts = hcounts.to(torch.int32)
    table = torch.full((P, 2), -1, dtype=torch.int32)
    table[:, 1] = 0
    table[hkeys.long(), 0] = (torch.cumsum(hcounts, 0) - hcounts).to(torch.int32)
    table[hkeys.long(), 1] = hcounts
    return dict(sx=sx, sh=sh, table=table.contiguous(),
                cell_keys=cell_keys[horder].to(torch.int32).contiguous(),
                counts=counts[horder].contiguous(), starts=starts[horder].contiguous(),
                sort=order.to(torch.int32).contiguous(), cells_x=cx, P=P, qmin=qmin,
                hmax=float(hmax))


def _pairs(rows, cols):
    r = rows.numpy().astype(np.int64)
    c = cols.numpy().astype(np.int64)
    o = np.lexsort((c, r))
    return r[o], c[o]


def self_search(x, h, vmin, vmax, radius=1):
    """(rows, cols) of buildNeighborList, lexicographically sorted."""
    m = load()
    s = hash_structures(x, h, vmin, vmax)
    rows, cols = m.buildNeighborList(s["sx"], s["sh"], s["table"], s["cell_keys"], s["counts"],
                                     s["starts"], s["sort"], s["cells_x"], s["P"], radius)
    return _pairs(rows, cols)


def foreign_search(xq, hq, x, h, vmin, vmax, radius=1):
    """(rows, cols) of buildNeighborListUnsortedPerParticle: queries xq
    against the grid of x, lexicographically sorted."""
    import torch

    m = load()
    s = hash_structures(x, h, vmin, vmax)
    xqt = torch.from_numpy(np.ascontiguousarray(xq, np.float32))
    hqt = torch.from_numpy(np.ascontiguousarray(hq, np.float32))
    rows, cols = m.buildNeighborListUnsortedPerParticle(
        xqt, hqt, s["sx"], s["sh"], s["table"], s["cell_keys"], s["counts"], s["starts"],
        s["sort"], s["cells_x"], s["P"], s["qmin"].contiguous(), s["hmax"], radius)
    return _pairs(rows, cols)
