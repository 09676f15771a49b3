"""sph_scan_exclusive_noalloc: the scan whose tile partials live in the
library's static device array, checked against sph_scan_exclusive with a
caller workspace (bit-exact, integer sums) on calls in a row, each smaller
than the one before, up to the largest size the static array admits."""
import numpy as np
import pytest
import torch

from pytorchsph_amd import _lib as L

pytestmark = pytest.mark.gpu

TILE = 2048
MAX_N = TILE * (65536 - 1)     # SPH_SCAN_NOALLOC_PARTIALS - 1 tiles


def scan_with_workspace(x: torch.Tensor, n: int) -> torch.Tensor:
    out = torch.empty(n + 1, dtype=torch.int32, device=x.device)
    lib = L.load()
    ws = torch.empty(int(lib.sph_scan_workspace_bytes(n)), dtype=torch.uint8, device=x.device)
    L.call("sph_scan_exclusive", L.ptr(x), n, L.ptr(out), L.ptr(ws), int(ws.numel()),
           L.stream_handle(None))
    return out


def scan_noalloc(x: torch.Tensor, n: int, out: torch.Tensor) -> None:
    L.call("sph_scan_exclusive_noalloc", L.ptr(x), n, L.ptr(out), L.stream_handle(None))


def test_noalloc_scan_matches_workspace_scan(cuda):
    gen = torch.Generator(device=cuda).manual_seed(7)
    # largest admitted size first, then smaller calls reusing the partials
    # the previous call left behind, then an unaligned (scalar-path) input
    for n, offset in ((MAX_N, 0), (3_000_001, 0), (100_000, 0), (4097, 1), (1, 0), (0, 0)):
        base = torch.randint(0, 4, (n + offset + 1,), dtype=torch.int32, device=cuda,
                             generator=gen)
        x = base[offset:offset + n] if n else base[:0]
        out = torch.full((n + 1,), -7, dtype=torch.int32, device=cuda)
        scan_noalloc(x, n, out)
        ref = scan_with_workspace(x, n)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), n
        if 0 < n <= 100_000:
            h = x.cpu().numpy().astype(np.int64)
            assert np.array_equal(out.cpu().numpy(), np.concatenate([[0], np.cumsum(h)])), n
        del base, x, out, ref


def test_noalloc_scan_rejects_sizes_beyond_the_static_array(cuda):
    """Checked on the host before anything is launched: no buffer of this
    size is touched."""
    x = torch.zeros(4, dtype=torch.int32, device=cuda)
    rc = L.load().sph_scan_exclusive_noalloc(L.ptr(x), MAX_N + 1, L.ptr(x), L.stream_handle(None))
    assert rc == 3                                     # SPH_E_SPACE
    assert b"exceeds the static partials" in L.load().sph_last_error()
    torch.cuda.synchronize()
