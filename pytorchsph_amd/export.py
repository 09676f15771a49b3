"""HDF5 export of a simulation in the reference's layout
(simulationBase.py:442-492 and 742-780), written through `h5` (the HDF5 C
library, no h5py).

File:
  attrs            the scheme and parameter summary of simulationBase.py:452-481
  boundaryInformation/   (staticBoundary) boundaryPosition, boundaryVelocity,
                   boundarySupport, boundaryRestDensity, boundaryArea,
                   boundaryNormals, boundaryBodyAssociation
  simulationExport/NNNNN/   one group per exported step, attrs time,
                   timestep, dt; datasets UID, fluidPosition / fluidVelocity
                   (the state at the start of the step), finalPosition /
                   finalVelocity (at its end), fluidArea, fluidDensity,
                   fluidSupport, fluidPressure, fluidAcceleration,
                   fluidGravity (point-source gravity), fluidDpdt
                   (continuum density), then each module's exportState
                   (boundary state when not static, fluidShifting).

The reference's export-interval update, `lastExport = (lastExport + 1 %
interval == 0)` (simulationBase.py:780), makes the interval test pass on
every step; every step is exported here too.
"""
from __future__ import annotations

import os
from datetime import datetime

import numpy as np
import torch

from . import h5


def _np(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


class SimulationExport:
    """The export file of one simulation; `sim` is an SPHSimulation whose
    modules are initialised."""

    def __init__(self, sim, path: str | None = None):
        c = sim.config
        if path is None:
            timestamp = datetime.now().strftime("%Y-%m-%d_%H-%M-%S")
            os.makedirs("./export/", exist_ok=True)
            path = "./export/%s - %s.hdf5" % (c["export"]["prefix"], timestamp)
        self.exportPath = path
        self.outFile = f = h5.File(path, "w")
        a = f.attrs
        a["simulationScheme"] = c["simulation"]["scheme"]
        a["boundaryScheme"] = c["simulation"]["boundaryScheme"]
        a["densityScheme"] = c["simulation"]["densityScheme"]
        a["EOSgamma"] = c["pressure"]["gamma"]
        a["fluidPressureTerm"] = c["pressure"]["fluidPressureTerm"]
        a["boundaryPressureTerm"] = c["pressure"]["boundaryPressureTerm"]
        a["shiftingScheme"] = c["shifting"]["scheme"]
        a["shiftingEnabled"] = c["shifting"]["enabled"]
        a["targetNeighbors"] = c["kernel"]["targetNeighbors"]
        a["defaultKernel"] = c["kernel"]["defaultKernel"]
        a["device"] = c["compute"]["device"]
        a["floatprecision"] = c["compute"]["floatprecision"]
        a["restDensity"] = c["fluid"]["restDensity"]
        a["c0"] = c["fluid"]["c0"]
        a["radius"] = c["particle"]["radius"]
        a["packing"] = c["particle"]["packing"]
        a["spacing"] = c["particle"]["spacing"]
        a["velocityDiffusionScheme"] = c["diffusion"]["velocityScheme"]
        a["densityDiffusionScheme"] = c["diffusion"]["densityScheme"]
        a["alphaDiffusion"] = c["diffusion"]["alpha"]
        a["deltaDiffusion"] = c["diffusion"]["delta"]
        a["kinematicDiffusion"] = c["diffusion"]["kinematic"]
        a["staticBoundary"] = c["export"]["staticBoundary"]
        a["initialDt"] = c["integration"]["dt"]
        a["integrationScheme"] = c["integration"]["scheme"]
        a["fixedDt"] = c["timestep"]["fixed"]
        g = c.get("gravity", {})
        if not g.get("pointSource", False):
            a["fluidGravity"] = g.get("magnitude", 9.81) * np.array(g.get("direction", [0, -1]))
        if c["export"]["staticBoundary"] and "boundaryPosition" in sim.perennialState:
            grp = f.create_group("boundaryInformation")
            p = sim.perennialState
            for key in ("boundaryPosition", "boundaryVelocity", "boundarySupport",
                        "boundaryRestDensity", "boundaryArea", "boundaryNormals",
                        "boundaryBodyAssociation"):
                grp.create_dataset(key, data=_np(p[key]))
            grp.close()
        self.dataGroup = f.create_group("simulationExport")
        self.exportCounter = 0

    def write_step(self, sim) -> None:
        """One group for the step just integrated (simulationBase.py:742-780)."""
        c = sim.config
        s, p = sim.simulationState, sim.perennialState
        grp = self.dataGroup.create_group("%05d" % self.exportCounter)
        self.exportCounter += 1
        mask = (s["ghostIndices"] == -1) if "ghostIndices" in s else s["UID"] > -1
        grp.attrs["time"] = float(s["time"])
        grp.attrs["timestep"] = int(s["timestep"])
        grp.attrs["dt"] = float(s["dt"])
        grp.create_dataset("UID", data=_np(s["UID"]))
        grp.create_dataset("fluidPosition", data=_np(p["fluidPosition"]))
        grp.create_dataset("fluidVelocity", data=_np(p["fluidVelocity"]))
        grp.create_dataset("finalPosition", data=_np(s["fluidPosition"]))
        grp.create_dataset("finalVelocity", data=_np(s["fluidVelocity"]))
        grp.create_dataset("fluidArea", data=_np(s["fluidArea"]))
        grp.create_dataset("fluidDensity", data=_np(s["fluidDensity"]))
        grp.create_dataset("fluidSupport", data=_np(s["fluidSupport"]))
        grp.create_dataset("fluidPressure", data=_np(s["fluidPressure"]))
        grp.create_dataset("fluidAcceleration", data=_np(s["fluidAcceleration"]))
        g = c.get("gravity", {})
        if g.get("pointSource", False):
            # written with the sign of simulationBase.py:762-768
            center = torch.tensor(g["center"], dtype=torch.float32,
                                  device=s["fluidPosition"].device)
            diff = s["fluidPosition"] - center
            dist = torch.linalg.norm(diff, axis=1)
            far = dist > 1e-7
            diff[far] = diff[far] / dist[far, None]
            val = g["magnitude"] * diff * dist[:, None] if g.get("potentialField", True) \
                else g["magnitude"] * diff
            grp.create_dataset("fluidGravity", data=_np(val))
        if c["simulation"]["densityScheme"] == "continuum":
            grp.create_dataset("fluidDpdt", data=_np(s["dpdt"]))
        for module in sim.modules:
            module.exportState(s, sim, grp, mask)
        grp.close()
        self.outFile.flush()

    def close(self) -> None:
        if self.dataGroup is not None:
            self.dataGroup.close()
            self.dataGroup = None
        self.outFile.close()


def write_job_result(path: str, scene, result: dict, steps: int) -> None:
    """A served job's result in the same layout: the file attributes the
    scene's constants give, boundaryInformation, and one group
    simulationExport/00000 with the job's final state (fluidPosition /
    fluidVelocity and finalPosition / finalVelocity both hold it; a served
    job keeps no start-of-step copy), its density, area, support, pressure
    and acceleration.  `result` is `Server`'s job.result (x, v, rho in the
    original particle order, status, and from SceneBatch.scene_state(extra)
    `p` and `acc`): fluidPressure is the DFSPH density solve's pressure
    (dfsph.py:307) or, for delta-SPH, the EOS of the last right-hand side's
    density; fluidAcceleration is the reference's last fluidAcceleration
    (simulationBase.py:569,701).  Without `p` the pressure is the EOS of the
    final density; without `acc` no fluidAcceleration is written.  Not
    written: fluidDpdt (delta-SPH continuity rate), which served jobs do not
    keep."""
    from .scenes import box_normals

    P = scene.phys
    dsph = scene.scheme == "dsph"
    with h5.File(path, "w") as f:
        a = f.attrs
        a["simulationScheme"] = "deltaSPH" if dsph else "dfsph"
        a["boundaryScheme"] = "solid" if scene.boundary == "solid" else "Akinci"
        a["densityScheme"] = "continuum" if dsph else "summation"
        a["EOSgamma"] = float(P["gamma"])
        a["fluidPressureTerm"] = "TaitEOS"
        a["boundaryPressureTerm"] = "PBSPH"
        a["shiftingScheme"] = "deltaPlus"
        a["shiftingEnabled"] = False
        a["targetNeighbors"] = 20
        a["defaultKernel"] = "wendland2"
        a["device"] = "cuda"
        a["floatprecision"] = "single"
        a["restDensity"] = float(scene.rho0)
        a["c0"] = float(P["c0"])
        a["radius"] = float(np.sqrt(scene.area / np.pi))
        a["packing"] = float(np.sqrt(scene.area) / scene.support)
        a["spacing"] = float(np.sqrt(scene.area) / scene.support)
        a["velocityDiffusionScheme"] = "deltaSPH"
        a["densityDiffusionScheme"] = P.get("diffusion", "MOG")
        a["alphaDiffusion"] = float(P["alpha"])
        a["deltaDiffusion"] = float(P["delta"])
        a["kinematicDiffusion"] = float(P["nu"])
        a["staticBoundary"] = True
        a["initialDt"] = float(P["dt"])
        a["integrationScheme"] = "RK4" if dsph else "semiImplicitEuler"
        a["fixedDt"] = True
        a["fluidGravity"] = np.array([P["gx"], P["gy"]], dtype=np.float64)
        a["jobStatus"] = int(result.get("status", 0))
        nb = scene.nb
        bx = np.asarray(scene.bx, np.float32)
        lo = np.asarray(scene.vmin, np.float64) + 2.0 * scene.support
        hi = np.asarray(scene.vmax, np.float64) - 2.0 * scene.support
        g = f.create_group("boundaryInformation")
        g.create_dataset("boundaryPosition", data=bx)
        g.create_dataset("boundaryVelocity", data=np.zeros((nb, 2), np.float32))
        g.create_dataset("boundarySupport", data=np.full(nb, scene.support, np.float32))
        g.create_dataset("boundaryRestDensity", data=np.full(nb, scene.rho0, np.float32))
        g.create_dataset("boundaryArea", data=np.asarray(scene.bvol, np.float32))
        g.create_dataset("boundaryNormals", data=box_normals(bx, lo, hi))
        g.create_dataset("boundaryBodyAssociation", data=np.zeros(nb, np.int64))
        g.close()
        d = f.create_group("simulationExport")
        grp = d.create_group("00000")
        grp.attrs["time"] = float(steps * P["dt"])
        grp.attrs["timestep"] = int(steps)
        grp.attrs["dt"] = float(P["dt"])
        x = np.asarray(result["x"], np.float32)
        v = np.asarray(result["v"], np.float32)
        rho = np.asarray(result["rho"], np.float32)
        n = len(x)
        if "p" in result:
            p = np.asarray(result["p"], np.float64)
        else:
            r = rho.astype(np.float64)
            p = scene.rho0 * P["c0"] ** 2 / P["gamma"] * (r ** P["gamma"] - 1.0)
        grp.create_dataset("UID", data=np.arange(n, dtype=np.int64))
        for key, val in (("fluidPosition", x), ("fluidVelocity", v), ("finalPosition", x),
                         ("finalVelocity", v)):
            grp.create_dataset(key, data=val)
        grp.create_dataset("fluidArea", data=np.full(n, scene.area, np.float32))
        grp.create_dataset("fluidDensity", data=rho)
        grp.create_dataset("fluidSupport", data=np.full(n, scene.support, np.float32))
        grp.create_dataset("fluidPressure", data=p.astype(np.float32))
        if "acc" in result:
            grp.create_dataset("fluidAcceleration", data=np.asarray(result["acc"], np.float32))
        grp.close()
        d.close()
