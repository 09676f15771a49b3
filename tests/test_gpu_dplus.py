"""delta-plus shifting on the device (deltaPlus.py:81-251, pressure.py:40-52)
through the module API, against the oracle fed with the same inputs:
renormalisation and lambda of fluid and boundary rows, the lambda gradient
and normal, the surface classes, the shift and the delta-plus pressure term;
then whole delta-plus steps with shifting against the oracle's step."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pytorchsph_amd.deltaSPH import deltaSPHSimulation
from tests.gpu_helpers import rel_err
from tests.test_gpu_module_api import base_config, scene_of

pytestmark = pytest.mark.gpu


def dplus_config(dx=0.02):
    c = base_config("dsph", dx)
    c["simulation"]["scheme"] = "deltaPlus"
    c["shifting"] = {"enabled": True, "scheme": "deltaPlus"}
    c["emitter"]["damBreak"]["velocity"] = [0.6, -0.4]
    return c


def jittered_sim(seed=0, dx=0.02):
    sim = deltaSPHSimulation(dplus_config(dx))
    sim.initializeSimulation()
    rng = np.random.default_rng(seed)
    s = sim.simulationState
    n = s["numParticles"]
    j = torch.from_numpy(rng.uniform(-0.15, 0.15, (n, 2)) * dx).float().cuda()
    s["fluidPosition"] += j
    s["fluidVelocity"] += torch.from_numpy(rng.uniform(-0.2, 0.2, (n, 2))).float().cuda()
    # continuum density off rest on both sides, so the pressure switch acts
    s["fluidDensity"] += torch.from_numpy(rng.uniform(-0.004, 0.004, n)).float().cuda()
    sim.perennialState = sim.saveState(copy=True)
    sim.simulationState = sim.setupSimulationState(sim.perennialState)
    return sim


def np64(t):
    return t.detach().cpu().numpy().astype(np.float64)


def orc_csr(c):
    rp = c.rowptr[: c.nrows + 1].cpu().numpy().astype(np.int64)
    return O.CSR(rp, c.col[: int(rp[-1])].cpu().numpy())


def surface_stage(sim):
    """Search, densities and detectSurface of one right-hand side."""
    s = sim.simulationState
    sim.neighborSearch.search(s, sim)
    sim.boundaryModule.boundaryFilterNeighborhoods(s, sim)
    sim.boundaryModule.boundaryNeighborhoodSearch(s, sim)
    sim.sphDensity.evaluate(s, sim)
    sim.boundaryModule.evalBoundaryDensity(s, sim)
    sim.shiftingModule.detectSurface(s, sim)
    torch.cuda.synchronize()
    bm = sim.boundaryModule
    P = dict(sim.physDict)
    f = dict(x=np64(s["fluidPosition"]), v=np64(s["fluidVelocity"]), rho=np64(s["fluidDensity"]),
             area=np64(s["fluidArea"]), vol=np64(s["fluidVolume"]))
    b = dict(x=np64(bm.boundaryPositions), rho=np64(bm.boundaryDensity),
             vol=np64(bm.boundaryVolume), p=np64(bm.pressure))
    lists = dict(ff=orc_csr(sim.neighborSearch.csr), fb=orc_csr(bm.fb), bf=orc_csr(bm.bf),
                 bb=orc_csr(bm.bb))
    return s, bm, P, f, b, lists


@pytest.fixture(scope="module")
def stage(cuda):
    sim = jittered_sim()
    return (sim,) + surface_stage(sim)


def test_boundary_pairs_match_brute_force(stage):
    sim, s, bm, P, f, b, lists = stage
    bx = bm.boundaryPositions.cpu().numpy()
    h = np.full(len(bx), np.float32(sim.config["particle"]["support"]), np.float32)
    ref = O.brute(bx, h, bx)
    assert np.array_equal(lists["bb"].rowptr, ref.rowptr)
    assert np.array_equal(lists["bb"].col, ref.col)


def test_lambda_fluid_and_boundary_rows(stage):
    sim, s, bm, P, f, b, lists = stage
    M, Lm, eig = O.fluid_norm(f, P, lists["ff"], b, lists["fb"])
    Mg = np64(s["normalizationMatrix"])
    assert np.abs(Mg - M).max() < 2e-5 * np.abs(M).max()
    assert np.abs(np64(s["eigVals"]) - eig).max() < 2e-5
    assert np.abs(np64(s["fluidLambda"]) - eig[:, 1]).max() < 2e-5
    well = eig[:, 1] > 0.05
    assert np.abs(np64(s["fluidL"])[well] - Lm[well]).max() < 1e-3 * np.abs(Lm[well]).max()
    vb = b["vol"] / b["rho"]
    Mb, Lb, eb = O.norm_rows(b["x"], P["h"], lists["bf"], f["x"], f["vol"], lists["bb"], b["x"],
                             vb)
    assert np.abs(np64(bm.normalizationMatrix) - Mb).max() < 2e-5 * np.abs(Mb).max()
    assert np.abs(np64(bm.eigVals) - eb).max() < 2e-5
    # a wetted wall row sees fluid on one side only, the bulk sees all around
    assert eig[:, 1].max() > 0.9 and eig[:, 1].min() < 0.6


def test_normal_matches_oracle(stage):
    sim, s, bm, P, f, b, lists = stage
    n = s["numParticles"]
    nb = bm.numPtcls
    m = min(n, nb)
    bvol_n = np.zeros(nb)
    bvol_n[:m] = f["area"][:m] / f["rho"][:m]
    grad, normal = O.dplus_normal(f, P, lists["ff"], np64(s["fluidL"]), np64(s["eigVals"]),
                                  b["x"], np64(bm.eigVals), bvol_n, lists["fb"])
    g = np64(s["lambdaGrad"])
    assert np.abs(g - grad).max() < 1e-4 * np.abs(grad).max()
    assert np.abs(np64(s["fluidNormal"]) - normal).max() < 1e-3
    # the free surface (top of the block) has outward normals
    top = f["x"][:, 1] > f["x"][:, 1].max() - 0.5 * math.sqrt(sim.config["particle"]["area"])
    assert (normal[top, 1] > 0.5).mean() > 0.8


def near_ties(cmax, lam):
    c6 = math.cos(math.pi / 6)
    return (np.abs(cmax - c6) < 1e-4) | (np.abs(lam - 0.6) < 1e-5) | (np.abs(lam - 0.5) < 1e-5)


def test_surface_classes_match_oracle(stage):
    sim, s, bm, P, f, b, lists = stage
    eig = np64(s["eigVals"])
    cmax, mask, surf = O.dplus_surface(f, P, lists["ff"], np64(s["fluidNormal"]), eig, b["x"],
                                       lists["fb"])
    got = np64(s["fluidSurfaceMask"])
    tie = near_ties(cmax, eig[:, 1])
    # a tie may flip its own class and those of its one ring
    ff = lists["ff"]
    rows, cols = ff.pairs()
    touched = tie.copy()
    np.logical_or.at(touched, rows, tie[cols])
    diff = got != surf
    assert not np.any(diff & ~touched), np.nonzero(diff & ~touched)[0][:10]
    assert touched.mean() < 0.02
    counts = {v: int((surf == v).sum()) for v in (0.0, 1.0, 2.0)}
    assert all(c > 0 for c in counts.values()), counts
    amin = np64(s["angleMin"])
    assert np.all((amin >= 0) & (amin <= math.pi))


def test_shift_matches_oracle(stage):
    sim, s, bm, P, f, b, lists = stage
    sm = sim.shiftingModule
    sm.shift(s, sim)
    torch.cuda.synchronize()
    raw, ref = O.dplus_shift(f, P, sm.packing, lists["ff"], np64(s["fluidNormal"]),
                             np64(s["fluidSurfaceMask"]), b, lists["fb"])
    got = np64(s["fluidUpdate"])
    h = P["h"]
    assert np.abs(got - ref).max() < 1e-4 * 0.05 * h
    assert np.all(np.linalg.norm(got, axis=1) <= 0.05 * h * (1 + 1e-6))
    assert np.abs(raw).max() > 0           # the shift is not trivially zero
    assert s["shiftAmount"] is s["fluidUpdate"]


def test_dplus_pressure_matches_oracle(stage):
    sim, s, bm, P, f, b, lists = stage
    s["fluidAcceleration"] = torch.zeros_like(s["fluidVelocity"])
    sim.pressureModule.computePressure(s, sim)
    sim.pressureModule.computePressureAcceleration(s, sim)
    torch.cuda.synchronize()
    b2 = dict(b, p=np64(bm.pressure))
    surf = np64(s["fluidSurfaceMask"])
    ref = O.dplus_pressure(f, P, lists["ff"], surf, b2, lists["fb"])
    got = np64(sim.pressureModule.pressureAccel)
    assert rel_err(got, ref) < 2e-4
    # the switch acts somewhere in this state (negative pressure in the bulk)
    p = O.eos(P, f["rho"])
    assert np.any((p < 0) & (surf > 1.5))


def test_dplus_steps_match_oracle(cuda):
    sim = jittered_sim(seed=3)
    sc = scene_of(sim, "dsph")
    so = O.SceneOracle(sc)
    packing = sim.shiftingModule.packing
    s0 = sim.simulationState
    x, v, r = np64(s0["fluidPosition"]), np64(s0["fluidVelocity"]), np64(s0["fluidDensity"])
    dx = math.sqrt(sim.config["particle"]["area"])
    for step in range(2):
        sim.integrate()
        torch.cuda.synchronize()
        x, v, r = so.dplus_step(x, v, r, packing)
        s = sim.simulationState
        gx, gv, gr = np64(s["fluidPosition"]), np64(s["fluidVelocity"]), np64(s["fluidDensity"])
        ex = np.abs(gx - x).max(axis=1)
        # a surface class decided by a near-tie may differ for a few rows
        assert np.quantile(ex, 0.999) < 1e-5 * dx / 0.02, (step, np.quantile(ex, 0.999))
        assert ex.max() < 0.1 * dx, (step, ex.max())
        assert np.quantile(np.abs(gr - r), 0.999) < 1e-5, step
        x, v, r = gx, gv, gr            # continue both from the device state


def test_export_of_a_dplus_run(cuda, tmp_path):
    """The module API's HDF5 export (simulationBase.py:442-492, 742-780) of a
    delta-plus run with shifting: one group per step, the end-of-step state,
    the shift of the step (deltaPlus.py:45-47) and the static boundary."""
    from pytorchsph_amd import h5

    c = dplus_config()
    c["export"] = {"active": True, "prefix": "dplus"}
    sim = deltaSPHSimulation(c)
    sim.exportPath = str(tmp_path / "run.hdf5")
    sim.initializeSimulation()
    for _ in range(2):
        sim.integrate()
    s = sim.simulationState
    sim.finalize()
    with h5.File(tmp_path / "run.hdf5", "r") as f:
        steps = f["simulationExport"]
        assert steps.keys() == ["00000", "00001"]
        g = steps["00001"]
        assert g.attrs["timestep"] == 2
        assert np.array_equal(g["finalPosition"][()], s["fluidPosition"].cpu().numpy())
        assert np.array_equal(g["fluidShifting"][()], s["fluidUpdate"].cpu().numpy())
        assert np.abs(g["fluidShifting"][()]).max() > 0
        bi = f["boundaryInformation"]
        n = bi["boundaryNormals"][()]
        assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
        assert f.attrs["shiftingEnabled"] == np.True_
        assert f.attrs["simulationScheme"] == "deltaPlus"


# ------------------------------------------------------ batched engine form
def dplus_scenes():
    from pytorchsph_amd import scenes as S

    out = []
    for k, (dx, block) in enumerate(((0.02, (0.6, 0.5)), (0.025, (0.5, 0.4)),
                                     (0.03, (0.45, 0.6)))):
        sc = S.dam_break(dx=dx, block=block, scheme="dsph", seed=20 + k)
        rng = np.random.default_rng(k)
        sc.v = rng.uniform(-0.3, 0.3, sc.x.shape).astype(np.float32)
        out.append(S.dplus_variant(sc, shifting=True))
    return out


def test_engine_dplus_steps_match_oracle(cuda):
    """The batched engine's delta-plus step (detection and pressure switch in
    every stage, the shift after the step) against the oracle's."""
    from pytorchsph_amd.engine import SceneBatch
    from pytorchsph_amd.scenes import lattice_calibration

    sc = dplus_scenes()[0]
    b = SceneBatch([sc])
    so = O.SceneOracle(sc)
    packing = math.sqrt(sc.area / lattice_calibration()[0]) / sc.support
    x, v, r = sc.x.astype(np.float64), sc.v.astype(np.float64), np.ones(sc.n)
    dx = math.sqrt(sc.area)
    for step in range(3):
        b.step()
        torch.cuda.synchronize()
        x, v, r = so.dplus_step(x, v, r, packing)
        got = b.scene_state(0)
        ex = np.abs(got["x"] - x).max(axis=1)
        assert np.quantile(ex, 0.999) < 1e-5 * dx / 0.02, (step, np.quantile(ex, 0.999))
        assert ex.max() < 0.1 * dx, (step, ex.max())
        assert np.quantile(np.abs(got["rho"] - r), 0.999) < 1e-5, step
        x, v, r = (got[k].astype(np.float64) for k in ("x", "v", "rho"))
    st, _ = b.check()
    assert np.all(st == 0)
    assert np.abs(b.shift.cpu().numpy()).max() > 0
    surf = b.smask.cpu().numpy()
    assert {0.0, 1.0, 2.0} <= set(np.unique(surf))


def test_engine_dplus_alone_equals_in_mix_and_graph(cuda):
    """Batched delta-plus scenes step bit for bit as they do alone, and a
    captured step replays the eager one; plain delta-SPH scenes in the same
    mixed batch run in their own part."""
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd.engine import MixedBatch, SceneBatch

    scenes = dplus_scenes()
    plain = S.dam_break(dx=0.03, block=(0.5, 0.5), scheme="dsph", seed=9)
    mix = MixedBatch(scenes + [plain])
    assert len(mix.parts) == 2
    alone = [SceneBatch([s]) for s in scenes]
    graph = SceneBatch(scenes)
    graph.capture(warm=False)
    for _ in range(3):
        mix.step()
        graph.step_graph()
        for a in alone:
            a.step()
    torch.cuda.synchronize()
    for k, a in enumerate(alone):
        ref = a.scene_state(0)
        for got in (mix.scene_state(k), graph.scene_state(k)):
            for key in ("x", "v", "rho"):
                assert np.array_equal(got[key], ref[key]), (k, key)


def test_switched_packed_gather_equals_dplus_pressure(stage):
    """sph_dsph_gather_packed_dplus with the pressure terms alone gives the
    acceleration of sph_dplus_pressure (the same switch, summed in another
    order)."""
    import ctypes as C

    from pytorchsph_amd import _lib as L
    from pytorchsph_amd.modules._ctx import csr_c, fluid

    sim, s, bm, P, f, b, lists = stage
    sim.pressureModule.computePressure(s, sim)
    n = s["numParticles"]
    dev = s["fluidPosition"].device
    F = fluid(sim, s, s.get("fluidVolume"))
    B = bm.boundary_struct()
    frec = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    brec = torch.zeros((bm.numPtcls, 8), dtype=torch.float32, device=dev)
    L.call("sph_dsph_pack", C.byref(F), L.ptr(sim.phys), C.byref(B), L.ptr(frec), L.ptr(brec),
           L.stream_handle())
    ns = sim.neighborSearch
    ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
    fb = bm.fb_struct()
    a1 = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    L.call("sph_dsph_gather_packed_dplus", L.ptr(frec), L.ptr(brec), L.ptr(sim.pscene(n)), n,
           L.ptr(sim.phys), C.byref(ff), C.byref(fb), None, L.T_PRESS | L.T_PRESS_B, None,
           L.ptr(s["fluidSurfaceMask"]), None, L.ptr(a1), L.stream_handle())
    a2 = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    L.call("sph_dplus_pressure", C.byref(F), L.ptr(sim.phys), C.byref(ff), C.byref(B),
           C.byref(fb), L.ptr(s["fluidSurfaceMask"]), L.T_PRESS | L.T_PRESS_B, L.ptr(a2),
           L.stream_handle())
    torch.cuda.synchronize()
    assert rel_err(a1.cpu().numpy(), a2.cpu().numpy()) < 1e-5


def test_fused_detect_equals_normal_then_surface(stage):
    """sph_dplus_detect (normal and angle test in one pass) gives the normal,
    angles and classes of sph_dplus_normal + sph_dplus_surface, bit for bit."""
    import ctypes as C

    from pytorchsph_amd import _lib as L
    from pytorchsph_amd.modules._ctx import csr_c, fluid

    sim, s, bm, P, f, b, lists = stage
    n = s["numParticles"]
    dev = s["fluidPosition"].device
    F = fluid(sim, s, s.get("fluidVolume"))
    ns = sim.neighborSearch
    ff = csr_c(ns.csr.rowptr, ns.csr.col, n)
    fb = bm.fb_struct()
    bvol = sim.shiftingModule._boundary_volumes_quirk(s, bm)
    out = {k: torch.zeros(n if k != "normal" else (n, 2), dtype=torch.float32, device=dev)
           for k in ("normal", "tmp", "amin", "amax", "surf")}
    L.call("sph_dplus_detect", C.byref(F), L.ptr(sim.phys), C.byref(ff), L.ptr(s["fluidL"]),
           L.ptr(s["eigVals"]), L.ptr(bm.boundaryPositions), L.ptr(bm.eigVals), L.ptr(bvol),
           C.byref(fb), None, L.ptr(out["normal"]), L.ptr(out["tmp"]), L.ptr(out["amin"]),
           L.ptr(out["amax"]), L.ptr(out["surf"]), L.stream_handle())
    sm = sim.shiftingModule
    sm.computeFluidNormal(s, sim)
    sm.detectFluidSurface(s, sim)
    torch.cuda.synchronize()
    assert torch.equal(out["normal"], s["fluidNormal"])
    assert torch.equal(out["surf"], s["fluidSurfaceMask"])
    assert torch.equal(out["amin"], s["angleMin"]) and torch.equal(out["amax"], s["angleMax"])
