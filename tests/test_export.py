"""HDF5 export in the reference's layout (simulationBase.py:442-492,
742-780) through the HDF5 C library: the writer's files read back through
this package's reader and through the HDF5 tools' own `h5dump`, an
independent reader of the format."""
import shutil
import subprocess
import types

import numpy as np
import pytest
import torch

from pytorchsph_amd import h5
from pytorchsph_amd.export import SimulationExport
from pytorchsph_amd.simulationBase import SPHSimulation

try:
    h5.lib()
    HAVE_H5 = True
except RuntimeError:
    HAVE_H5 = False

pytestmark = pytest.mark.skipif(not HAVE_H5, reason="no HDF5 C library in this image")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if shutil.which("/opt/conda/bin/h5dump") else None)

# simulationBase.py:452-481 (fluidGravity for constant gravity) and :748-772
FILE_ATTRS = {
    "simulationScheme", "boundaryScheme", "densityScheme", "EOSgamma", "fluidPressureTerm",
    "boundaryPressureTerm", "shiftingScheme", "shiftingEnabled", "targetNeighbors",
    "defaultKernel", "device", "floatprecision", "restDensity", "c0", "radius", "packing",
    "spacing", "velocityDiffusionScheme", "densityDiffusionScheme", "alphaDiffusion",
    "deltaDiffusion", "kinematicDiffusion", "staticBoundary", "initialDt",
    "integrationScheme", "fixedDt", "fluidGravity"}
BOUNDARY_SETS = {"boundaryPosition", "boundaryVelocity", "boundarySupport",
                 "boundaryRestDensity", "boundaryArea", "boundaryNormals",
                 "boundaryBodyAssociation"}
STEP_SETS = {"UID", "fluidPosition", "fluidVelocity", "finalPosition", "finalVelocity",
             "fluidArea", "fluidDensity", "fluidSupport", "fluidPressure", "fluidAcceleration",
             "fluidDpdt"}


def fake_sim(n=37, nb=11, seed=0):
    """A parsed config and CPU state dictionaries of the shapes a simulation
    holds (the writer only reads them)."""
    cfg = {"particle": {"radius": 0.01}, "simulation": {"scheme": "deltaSPH",
                                                        "densityScheme": "continuum"},
           "export": {"active": True, "staticBoundary": True}}
    sim = types.SimpleNamespace()
    for p in SPHSimulation.getBasicParameters(sim):
        p.parseConfig(cfg)
    cfg["particle"].update(area=np.pi * 1e-4, support=np.float32(0.0447), packing=0.39,
                           spacing=0.39)
    cfg["fluid"]["c0"] = 24.26
    rng = np.random.default_rng(seed)
    t = lambda *s: torch.from_numpy(rng.standard_normal(s).astype(np.float32))  # noqa: E731
    p = {"fluidPosition": t(n, 2), "fluidVelocity": t(n, 2),
         "boundaryPosition": t(nb, 2), "boundaryVelocity": t(nb, 2), "boundarySupport": t(nb),
         "boundaryRestDensity": t(nb), "boundaryArea": t(nb), "boundaryNormals": t(nb, 2),
         "boundaryBodyAssociation": torch.zeros(nb, dtype=torch.int64)}
    s = {"UID": torch.arange(n, dtype=torch.int64), "fluidPosition": t(n, 2),
         "fluidVelocity": t(n, 2), "fluidArea": t(n), "fluidDensity": t(n),
         "fluidSupport": t(n), "fluidPressure": t(n), "fluidAcceleration": t(n, 2),
         "dpdt": t(n), "ghostIndices": torch.full((n,), -1, dtype=torch.int64),
         "time": 0.002, "timestep": 3, "dt": 0.001}
    sim.config, sim.perennialState, sim.simulationState, sim.modules = cfg, p, s, []
    return sim


def test_layout_and_values_round_trip(tmp_path):
    sim = fake_sim()
    path = tmp_path / "run.hdf5"
    ex = SimulationExport(sim, str(path))
    ex.write_step(sim)
    sim.simulationState["timestep"] = 4
    ex.write_step(sim)
    ex.close()
    with h5.File(path, "r") as f:
        assert set(f.keys()) == {"boundaryInformation", "simulationExport"}
        assert set(f.attrs.keys()) == FILE_ATTRS
        assert f.attrs["simulationScheme"] == "deltaSPH"
        assert f.attrs["shiftingEnabled"] == np.False_
        assert f.attrs["targetNeighbors"] == 20
        assert np.array_equal(f.attrs["fluidGravity"], [0.0, -9.81])
        b = f["boundaryInformation"]
        assert set(b.keys()) == BOUNDARY_SETS
        for k in BOUNDARY_SETS:
            assert np.array_equal(b[k][()], sim.perennialState[k].numpy()), k
        steps = f["simulationExport"]
        assert steps.keys() == ["00000", "00001"]
        g = steps["00001"]
        assert set(g.keys()) == STEP_SETS
        assert g.attrs["timestep"] == 4 and g.attrs["dt"] == 0.001
        s = sim.simulationState
        assert np.array_equal(g["finalPosition"][()], s["fluidPosition"].numpy())
        assert np.array_equal(g["fluidPosition"][()], sim.perennialState["fluidPosition"].numpy())
        assert np.array_equal(g["fluidDpdt"][()], s["dpdt"].numpy())
        assert g["UID"][()].dtype == np.int64
        assert g["fluidAcceleration"][()].shape == (37, 2)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_h5dump_reads_the_file(tmp_path):
    sim = fake_sim(n=5, nb=3)
    path = tmp_path / "run.hdf5"
    ex = SimulationExport(sim, str(path))
    ex.write_step(sim)
    ex.close()
    hdr = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True)
    out = hdr.stdout
    assert 'GROUP "simulationExport"' in out and 'GROUP "00000"' in out
    assert 'DATASET "finalPosition"' in out and "SIMPLE { ( 5, 2 ) / ( 5, 2 ) }" in out
    assert 'ATTRIBUTE "simulationScheme"' in out and "H5T_VARIABLE" in out
    data = subprocess.run([H5DUMP, "-d", "/simulationExport/00000/UID", "-y", "-w", "0",
                           str(path)], capture_output=True, text=True, check=True).stdout
    assert "0, 1, 2, 3, 4" in data.replace("\n", " ")
    attr = subprocess.run([H5DUMP, "-a", "/simulationScheme", str(path)], capture_output=True,
                          text=True, check=True).stdout
    assert '"deltaSPH"' in attr


def test_point_source_gravity_and_module_hook(tmp_path):
    sim = fake_sim(n=6, nb=2)
    sim.config["gravity"] = {"magnitude": 2.0, "pointSource": True, "potentialField": False,
                             "center": [0.0, 0.0], "direction": [0, -1]}
    sim.config["export"]["staticBoundary"] = False

    class Hook:
        def exportState(self, state, s, grp, mask):
            grp.create_dataset("fluidShifting", data=np.ones((6, 2), np.float32))

    sim.modules = [Hook()]
    path = tmp_path / "run.hdf5"
    ex = SimulationExport(sim, str(path))
    ex.write_step(sim)
    ex.close()
    with h5.File(path, "r") as f:
        assert "fluidGravity" not in f.attrs and "boundaryInformation" not in f
        g = f["simulationExport"]["00000"]
        x = sim.simulationState["fluidPosition"].numpy().astype(np.float64)
        ref = 2.0 * x / np.linalg.norm(x, axis=1, keepdims=True)
        assert np.allclose(g["fluidGravity"][()], ref, atol=1e-6)
        assert np.array_equal(g["fluidShifting"][()], np.ones((6, 2), np.float32))


def test_attribute_types(tmp_path):
    path = tmp_path / "a.hdf5"
    with h5.File(path, "w") as f:
        f.attrs["s"] = "text"
        f.attrs["b"] = True
        f.attrs["i"] = 7
        f.attrs["f"] = 0.5
        f.attrs["f32"] = np.float32(0.25)
        f.attrs["arr"] = np.arange(3, dtype=np.int32)
        f.create_dataset("flags", data=np.array([True, False]))
        f.create_dataset("empty", data=np.zeros((0, 2), np.float32))
    with h5.File(path, "r") as f:
        assert f.attrs["s"] == "text" and f.attrs["b"] == np.True_
        assert f.attrs["i"] == 7 and f.attrs["i"].dtype == np.int64
        assert f.attrs["f"] == 0.5 and f.attrs["f32"].dtype == np.float32
        assert np.array_equal(f.attrs["arr"], [0, 1, 2])
        assert np.array_equal(f["flags"][()], [True, False])
        assert f["empty"][()].shape == (0, 2)


def test_served_job_result_layout(tmp_path):
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd.export import write_job_result

    sc = S.dam_break(dx=0.05, block=(0.5, 0.5), tank=(1.0, 1.0), scheme="dsph", seed=1)
    rng = np.random.default_rng(0)
    res = dict(x=sc.x + 0.01, v=rng.standard_normal(sc.x.shape).astype(np.float32),
               rho=np.full(sc.n, 1.001, np.float32), status=0)
    path = tmp_path / "job.hdf5"
    write_job_result(str(path), sc, res, steps=17)
    with h5.File(path, "r") as f:
        assert FILE_ATTRS <= set(f.attrs.keys())
        assert set(f["boundaryInformation"].keys()) == BOUNDARY_SETS
        g = f["simulationExport"]["00000"]
        assert g.attrs["timestep"] == 17
        assert np.array_equal(g["finalPosition"][()], res["x"])
        p = g["fluidPressure"][()]
        r = float(np.float32(1.001))      # the stored fp32 density
        ref = 1000.0 * sc.phys["c0"] ** 2 / 7.0 * (r ** 7 - 1)
        assert np.allclose(p, ref, rtol=1e-5)
        assert "fluidAcceleration" not in g     # none in the result: not written
        nrm = f["boundaryInformation"]["boundaryNormals"][()]
        bx = f["boundaryInformation"]["boundaryPosition"][()]
        # outward: walls below the tank point down, walls left of it point left
        assert np.all(nrm[bx[:, 1] < -0.05, 1] < 0) and np.all(nrm[bx[:, 0] < -0.05, 0] < 0)


def test_served_job_result_writes_carried_pressure_and_acceleration(tmp_path):
    """A DFSPH job's result carries the solver pressure and the last
    acceleration (SceneBatch.scene_state(extra=True)); both are written as
    they are, not recomputed from the density."""
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd.export import write_job_result

    sc = S.dam_break(dx=0.05, block=(0.5, 0.5), tank=(1.0, 1.0), scheme="dfsph", seed=1)
    rng = np.random.default_rng(1)
    res = dict(x=sc.x, v=sc.v, rho=np.ones(sc.n, np.float32), status=0,
               p=rng.uniform(0, 50, sc.n).astype(np.float32),
               acc=rng.standard_normal((sc.n, 2)).astype(np.float32))
    path = tmp_path / "job.hdf5"
    write_job_result(str(path), sc, res, steps=3)
    with h5.File(path, "r") as f:
        g = f["simulationExport"]["00000"]
        assert np.array_equal(g["fluidPressure"][()], res["p"])
        assert np.array_equal(g["fluidAcceleration"][()], res["acc"])
