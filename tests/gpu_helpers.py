"""Shared helpers of the GPU parity tests."""
import numpy as np
import torch

from pytorchsph_amd.search import NeighborGrid, grid_params


def csr_numpy(rowptr, col, nrows):
    rp = rowptr[: nrows + 1].detach().cpu().numpy().astype(np.int64)
    c = col[: max(int(rp[-1]), 0)].detach().cpu().numpy().astype(np.int64)
    return rp, c


def sorted_rows(rp, col):
    out = col.copy()
    for r in range(len(rp) - 1):
        a, b = rp[r], rp[r + 1]
        if b - a > 1:
            out[a:b] = np.sort(out[a:b])
    return out


def assert_same_pairs(rp_gpu, col_gpu, orc_csr, sort=True):
    rp_o = orc_csr.rowptr
    assert np.array_equal(rp_gpu, rp_o), "row lengths differ"
    c = sorted_rows(rp_gpu, col_gpu) if sort else col_gpu
    assert np.array_equal(c, orc_csr.col.astype(np.int64)), "columns differ"


def gpu_search(x, h, vmin, vmax, hmax, queries=None, radius=1, ordered=True):
    """Self search (queries None) or foreign queries (xq, hq) on one scene;
    ordered False: the grid by the radix-sort build."""
    n = len(x)
    dev = torch.device("cuda")
    g = NeighborGrid([dict(vmin=vmin, vmax=vmax, hmax=hmax, p_begin=0, p_count=n)], n, dev)
    xt = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    ps = torch.zeros(n, dtype=torch.int32, device=dev)
    g.build(xt, ps, ordered=ordered)
    if queries is None:
        ht = torch.from_numpy(np.ascontiguousarray(h, np.float32)).to(dev)
        csr = g.search(xt, ht, ps, radius, self_sorted=True)
    else:
        xq, hq = queries
        xqt = torch.from_numpy(np.ascontiguousarray(xq, np.float32)).to(dev)
        hqt = torch.from_numpy(np.ascontiguousarray(hq, np.float32)).to(dev)
        qs = torch.zeros(len(xq), dtype=torch.int32, device=dev)
        csr = g.search(xqt, hqt, qs, radius, self_sorted=False)
    torch.cuda.synchronize()
    st = int(g.status.cpu()[0])
    rp, c = csr_numpy(csr.rowptr, csr.col, csr.nrows)
    return rp, c, st, g


def in_scene_order(b, t, k=0):
    """Fluid array t of scene k of SceneBatch b in the scene's original
    particle order (the batch keeps its state in cell order)."""
    a, e = int(b.p_begin[k]), int(b.p_begin[k] + b.n[k])
    inv = np.empty(e - a, dtype=np.int64)
    inv[b.uid[a:e].cpu().numpy() - a] = np.arange(e - a)
    return t[a:e].detach().cpu().numpy()[inv]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-30)
    return float(np.max(np.abs(a - b)) / scale)


def grid_of(scene):
    return grid_params(scene.vmin, scene.vmax, scene.support)


class tuned:
    """Sets launch-shape knobs for the duration of a with-block."""

    def __init__(self, **lanes):
        self.lanes = lanes
        self.old = None

    def __enter__(self):
        from pytorchsph_amd import _lib as L
        self.old = L.set_tuning(**self.lanes)
        return self

    def __exit__(self, *exc):
        from pytorchsph_amd import _lib as L
        L.set_tuning(**self.old)
        return False
