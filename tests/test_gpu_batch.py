"""Batched form: every scene's result is the same alone and in the mix, and
a captured hipGraph step reproduces the eager step."""
import numpy as np
import pytest
import torch

from pytorchsph_amd import scenes as S
from pytorchsph_amd.engine import SceneBatch

pytestmark = pytest.mark.gpu


def mix(scheme):
    return [S.dam_break(dx=0.02, scheme=scheme, seed=1),
            S.colliding_drops(dx=0.015, scheme=scheme, seed=2),
            S.dam_break(dx=0.03, block=(0.6, 0.9), scheme=scheme, seed=3)]


@pytest.mark.parametrize("scheme", ["dsph", "dfsph"])
def test_alone_equals_in_mix(cuda, scheme):
    scenes = mix(scheme)
    batch = SceneBatch(scenes)
    alone = [SceneBatch([s]) for s in scenes]
    for _ in range(3):
        batch.step()
        for a in alone:
            a.step()
    torch.cuda.synchronize()
    for k, a in enumerate(alone):
        got = batch.scene_state(k)
        ref = a.scene_state(0)
        if scheme == "dsph":
            # identical pair lists and summation order: bitwise equal
            for key in ("x", "v", "rho"):
                assert np.array_equal(got[key], ref[key]), key
        else:
            # the FP64 residual is summed with atomics, so iteration counts
            # (and results) agree to rounding
            assert np.abs(got["x"] - ref["x"]).max() < 1e-6
            assert np.abs(got["v"] - ref["v"]).max() < 1e-4
    if scheme == "dfsph":
        it_b = batch.dfsph_iterations()
        it_a = np.array([a.dfsph_iterations()[0] for a in alone])
        assert np.all(np.abs(it_b - it_a) <= 1)


@pytest.mark.parametrize("scheme", ["dsph", "dfsph"])
def test_graph_replay_matches_eager(cuda, scheme):
    scenes = mix(scheme)
    eager = SceneBatch(scenes)
    graph = SceneBatch(scenes)
    graph.capture()           # capture() runs one eager warm-up step
    eager.step()
    for _ in range(3):
        graph.step_graph()
        eager.step()
        torch.cuda.synchronize()
        # eager work between replays, then compare every replay
        for k in range(len(scenes)):
            g, e = graph.scene_state(k), eager.scene_state(k)
            if scheme == "dsph":
                assert np.array_equal(g["x"], e["x"]) and np.array_equal(g["v"], e["v"])
            else:
                assert np.abs(g["x"] - e["x"]).max() < 1e-6
    st, _ = graph.check()
    # the iteration cap may be reached on an unrelaxed lattice (the oracle
    # reaches it too); pair-list overflow, lost points or NaN may not occur
    from pytorchsph_amd import _lib as L
    assert np.all(st & (L.ST_OUTSIDE | L.ST_OVERFLOW | L.ST_NAN) == 0)


@pytest.mark.parametrize("scheme", ["dsph", "dfsph"])
def test_cold_capture_matches_eager(cuda, scheme):
    """capture(warm=False) advances nothing; replays from there follow the
    eager steps of a fresh batch (the first reorder is the identity)."""
    scenes = mix(scheme)
    eager = SceneBatch(scenes)
    graph = SceneBatch(scenes)
    graph.capture(warm=False)
    torch.cuda.synchronize()
    for k in range(len(scenes)):
        assert np.array_equal(graph.scene_state(k)["x"], scenes[k].x)
    for _ in range(3):
        graph.step_graph()
        eager.step()
        torch.cuda.synchronize()
        for k in range(len(scenes)):
            g, e = graph.scene_state(k), eager.scene_state(k)
            if scheme == "dsph":
                assert np.array_equal(g["x"], e["x"]) and np.array_equal(g["v"], e["v"])
            else:
                assert np.abs(g["x"] - e["x"]).max() < 1e-6


@pytest.mark.parametrize("reorder", [True, False])
def test_fused_rk4_step_matches_separate_kernels(cuda, reorder):
    """sph_dsph_gather_rk4 (gather + RK4 stage combine + next records in one
    launch) gives the state of the separate gather, integrate and pack
    kernels, bit for bit, over several steps."""
    scenes = mix("dsph")
    fused = SceneBatch(scenes, reorder=reorder, fuse_rk4=True)
    split = SceneBatch(scenes, reorder=reorder, fuse_rk4=False)
    for _ in range(4):
        fused.step()
        split.step()
    torch.cuda.synchronize()
    for k in range(len(scenes)):
        a, b = fused.scene_state(k), split.scene_state(k)
        for key in ("x", "v", "rho"):
            d = np.abs(a[key].astype(np.float64) - b[key])
            assert np.array_equal(a[key], b[key]), (k, key, int((d > 0).sum()), a[key].size,
                                                    float(d.max()), float(np.abs(b[key]).max()))


@pytest.mark.parametrize("n", [100_003, 100_000])
def test_permute_fields_matches_numpy(cuda, n):
    """sph_permute_fields moves every field by the same permutation, int64
    ids bit for bit, with fields of widths 1, 2 and 4, for an odd count and
    a multiple of four (the temporary's field blocks padded to 16 bytes)."""
    import ctypes as C

    from pytorchsph_amd import _lib as L

    rng = np.random.default_rng(5)
    perm = rng.permutation(n).astype(np.int32)
    a = rng.standard_normal((n, 2)).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    ids = rng.integers(-2**62, 2**62, size=n, dtype=np.int64)
    c = rng.standard_normal((n, 4)).astype(np.float32)
    dev = torch.device("cuda")
    ta, tb, tid, tc = (torch.from_numpy(t.copy()).to(dev) for t in (a, b, ids, c))
    tmp = torch.empty(n * 9 + 16, dtype=torch.float32, device=dev)
    fl = L.SphFields()
    for q, (t, w) in enumerate(((ta, 2), (tb, 1), (tid, 2), (tc, 4))):
        fl.ptr[q] = L.ptr(t)
        fl.width[q] = w
    fl.nfields = 4
    L.call("sph_permute_fields", C.byref(fl), L.ptr(torch.from_numpy(perm).to(dev)), n,
           L.ptr(tmp), L.stream_handle())
    torch.cuda.synchronize()
    assert np.array_equal(ta.cpu().numpy(), a[perm])
    assert np.array_equal(tb.cpu().numpy(), b[perm])
    assert np.array_equal(tid.cpu().numpy(), ids[perm])
    assert np.array_equal(tc.cpu().numpy(), c[perm])


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_finish_visc_matches_finish_then_gather(cuda, lanes):
    """sph_dfsph_finish_visc (the final pressure acceleration and the
    velocity-diffusion / laminar gather over one walk of the fluid list)
    gives acc, v* and a* of sph_dfsph_finish followed by
    sph_dsph_gather_packed(SPH_T_VISC | SPH_T_LAM | SPH_T_ACCUM), bit for bit
    when the solve and the gathers take the same lanes per row."""
    import ctypes as C

    from pytorchsph_amd import _lib as L

    old = L.set_tuning(solve=lanes, gather=lanes)
    try:
        b = SceneBatch(mix("dfsph"))
        b.step()                         # a flowed state with pressures
        b._dfsph_pre()
        b._dfsph_loop()
        torch.cuda.synchronize()
        keep = {k: b.df[k].clone() for k in ("vpred", "apred", "p")}
        acc0 = b.acc.clone()
        ff = b.ff.csr().c()
        fb = b.fb.csr().c() if b.NB else None
        fbp = C.byref(fb) if fb is not None else None
        F = b.fluid(b.x, b.v, b.rho)
        Fv = b.fluid(b.x, b.v, b.rho, b.vol)
        D = b._dfsph_struct()
        sh = L.stream_handle()
        L.call("sph_dsph_pack", C.byref(Fv), L.ptr(b.phys), None, L.ptr(b.frec), None, sh)
        L.call("sph_dfsph_finish", C.byref(F), L.ptr(b.phys), C.byref(ff), fbp, C.byref(D),
               L.ptr(b.acc), sh)
        L.call("sph_dsph_gather_packed", L.ptr(b.frec), None, L.ptr(b.pscene), b.NP,
               L.ptr(b.phys), C.byref(ff), None, None, L.T_VISC | L.T_LAM | L.T_ACCUM, None,
               None, L.ptr(b.acc), sh)
        torch.cuda.synchronize()
        split = {k: b.df[k].clone() for k in keep}
        split_acc = b.acc.clone()
        for k, v in keep.items():
            b.df[k].copy_(v)
        b.acc.copy_(acc0)
        L.call("sph_dfsph_finish_visc", C.byref(F), L.ptr(b.phys), C.byref(ff), fbp,
               C.byref(D), L.ptr(b.frec), L.T_VISC | L.T_LAM, L.ptr(b.acc), sh)
        torch.cuda.synchronize()
        assert torch.equal(b.acc, split_acc)
        assert not torch.equal(b.acc, acc0)
        for k in keep:
            assert torch.equal(b.df[k], split[k]), k
    finally:
        L.set_tuning(**old)
