"""SPHSimulation host driver with the reference's API
(src/simulationBase.py:21-782): config parsing with the same namespaces and
defaults, the simulationState dict, saveState / setupSimulationState and the
integrators (explicitEuler, semiImplicitEuler, PECE, PEC, RK4).

Scene set-up differs from the reference, which samples SDF polygons and
optimises the packing with torch_cluster: emitters are filled on a lattice
of spacing sqrt(area) (box, or the inscribed disc for shape = 'sphere') and
the domain walls are generated directly (see scenes.py).
"""
from __future__ import annotations

import numpy as np
import torch

from .export import SimulationExport
from .kernels import getKernelFunctions
from .modules._ctx import phys_tensor
from .parameter import Parameter


class SPHSimulation:
    def getBasicParameters(self):
        P = Parameter
        return [
            P("particle", "radius", "float", 0.014426521330548324),
            P("simulation", "scheme", "string", "dfsph"),
            P("simulation", "mode", "string", "default"),
            P("simulation", "verbose", "bool", False),
            P("simulation", "boundaryScheme", "string", "solid"),
            P("simulation", "bodyForces", "bool", True),
            P("simulation", "densityScheme", "string", "summation"),
            P("akinciBoundary", "gamma", "float", 0.7),
            P("pressure", "gamma", "float", 7.0),
            P("pressure", "kappa", "float", 1.3),
            P("pressure", "fluidPressureTerm", "str", "TaitEOS"),
            P("pressure", "boundaryPressureTerm", "str", "PBSPH"),
            P("shifting", "scheme", "string", "deltaPlus"),
            P("shifting", "enabled", "bool", False),
            P("kernel", "targetNeighbors", "int", 20),
            P("kernel", "defaultKernel", "string", "wendland2"),
            P("compute", "maxNeighbors", "int", 256),
            P("compute", "device", "string", "cuda"),
            P("compute", "floatprecision", "string", "single"),
            P("fluid", "restDensity", "float", 1000),
            P("fluid", "gravity", "float array", [0, 0]),
            P("fluid", "backgroundPressure", "bool", False),
            P("fluid", "c0", "float", -1),
            P("integration", "dt", "float", 0.002),
            P("integration", "scheme", "string", "RK4"),
            P("diffusion", "velocityScheme", "string", "deltaSPH"),
            P("diffusion", "densityScheme", "string", "deltaSPH"),
            P("diffusion", "alpha", "float", 0.01),
            P("diffusion", "delta", "float", 0.1),
            P("diffusion", "kinematic", "float", 0.01),
            P("diffusion", "boundaryDiffusion", "bool", True),
            P("domain", "min", "float array", [-1, -1]),
            P("domain", "max", "float array", [1, 1]),
            P("domain", "adjustDomain", "bool", False),
            P("domain", "adjustParticle", "bool", False),
            P("export", "active", "bool", False),
            P("export", "prefix", "string", "unnamed"),
            P("export", "staticBoundary", "bool", True),
            P("export", "interval", "int", -1),
            P("periodicBC", "periodicX", "bool", False),
            P("periodicBC", "periodicY", "bool", False),
            P("periodicBC", "buffer", "float", 2),
            P("timestep", "fixed", "bool", True),
        ]

    def __init__(self, config):
        for param in self.getBasicParameters():
            param.parseConfig(config)
        self.config = config
        self.verbose = config["simulation"]["verbose"]
        self.kernel, self.kernelGrad = getKernelFunctions(config["kernel"]["defaultKernel"])
        if config["compute"]["floatprecision"] != "single":
            raise NotImplementedError("the device path is fp32")
        config["compute"]["precision"] = torch.float32
        self.dtype = torch.float32
        self.device = config["compute"]["device"]
        if config["periodicBC"]["periodicX"] or config["periodicBC"]["periodicY"]:
            raise NotImplementedError("periodic boundaries are outside the hot path")
        if config["shifting"]["enabled"] and config["shifting"]["scheme"] != "deltaPlus":
            raise NotImplementedError("only the deltaPlus shifting scheme is on the hot path")
        # simulationBase.py:848-855
        p = config["particle"]
        p["area"] = np.pi * p["radius"] ** 2
        p["support"] = np.single(np.sqrt(p["area"] / np.pi * config["kernel"]["targetNeighbors"]))
        p["packing"] = float(np.sqrt(p["area"]) / p["support"])
        p["spacing"] = p["packing"]
        # simulationBase.py:968-969
        d = config["domain"]
        buf = float(p["support"]) * config["periodicBC"]["buffer"]
        d["virtualMin"] = np.array(d["min"], dtype=np.float64) - buf
        d["virtualMax"] = np.array(d["max"], dtype=np.float64) + buf
        for e in config.get("emitter", {}).values():
            e.setdefault("restDensity", config["fluid"]["restDensity"])
            e.setdefault("shape", "box")
            e.setdefault("velocity", [0.0, 0.0])
        self._pscene = {}

    # ------------------------------------------------------------ helpers
    def pscene(self, n):
        t = self._pscene.get(n)
        if t is None:
            t = torch.zeros(n, dtype=torch.int32, device=self.device)
            self._pscene[n] = t
        return t

    def sync(self, tensor):
        return

    def _emit(self, e):
        dx = float(np.sqrt(self.config["particle"]["area"]))
        lo, hi = np.array(e["min"], dtype=np.float64), np.array(e["max"], dtype=np.float64)
        nx, ny = int(np.floor((hi[0] - lo[0]) / dx + 1e-9)), int(np.floor((hi[1] - lo[1]) / dx + 1e-9))
        I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
        pts = np.stack([lo[0] + (I.ravel() + 0.5) * dx, lo[1] + (J.ravel() + 0.5) * dx], 1)
        if e["shape"] == "sphere":
            c = (lo + hi) / 2
            r = min(hi[0] - lo[0], hi[1] - lo[1]) / 2
            pts = pts[((pts - c) ** 2).sum(1) <= r * r]
        return pts

    # ------------------------------------------------------------- set-up
    def initializeSimulation(self):
        c = self.config
        pos, vel, area, rest = [], [], [], []
        for e in c.get("emitter", {}).values():
            pts = self._emit(e)
            pos.append(pts)
            vel.append(np.tile(e["velocity"], (len(pts), 1)))
            area.append(np.full(len(pts), c["particle"]["area"]))
            rest.append(np.full(len(pts), e["restDensity"]))
        dev, f32 = self.device, torch.float32
        s = {}
        s["fluidPosition"] = torch.tensor(np.concatenate(pos), dtype=f32, device=dev)
        n = s["fluidPosition"].shape[0]
        s["UID"] = torch.arange(n, dtype=torch.int64, device=dev)
        s["ghostIndices"] = torch.full((n,), -1, dtype=torch.int64, device=dev)
        s["fluidDensity"] = torch.ones(n, dtype=f32, device=dev)
        s["fluidArea"] = torch.tensor(np.concatenate(area), dtype=f32, device=dev)
        s["fluidSupport"] = torch.full((n,), float(c["particle"]["support"]), dtype=f32, device=dev)
        s["fluidVelocity"] = torch.tensor(np.concatenate(vel), dtype=f32, device=dev)
        s["fluidAcceleration"] = torch.zeros((n, 2), dtype=f32, device=dev)
        s["fluidPressure"] = torch.zeros(n, dtype=f32, device=dev)
        s["fluidRestDensity"] = torch.tensor(np.concatenate(rest), dtype=f32, device=dev)
        s["dpdt"] = torch.zeros(n, dtype=f32, device=dev)
        s["numParticles"] = n
        s["realParticles"] = n
        s["time"] = 0.0
        s["timestep"] = 0
        s["dt"] = c["integration"]["dt"]
        self.simulationState = s
        for module in self.modules:
            module.initialize(self.config, self.simulationState)
        self.phys, self.physDict = phys_tensor(self.config, self.device)
        self.perennialState = self.saveState(copy=False)
        self.simulationState = self.setupSimulationState(self.perennialState)
        self.exporter = None
        if c["export"]["active"]:
            # simulationBase.py:442-492
            self.exporter = SimulationExport(self, getattr(self, "exportPath", None))
            self.exportPath = self.exporter.exportPath
            self.outFile = self.exporter.outFile
            self.dataGroup = self.exporter.dataGroup

    def finalize(self):
        """Closes the export file, if any."""
        if getattr(self, "exporter", None) is not None:
            self.exporter.close()
            self.exporter = None

    def resetState(self):
        for module in self.modules:
            module.resetState(self.simulationState)
        self.simulationState.pop("fluidAcceleration", None)

    def saveState(self, copy=False):
        """simulationBase.py:505-533."""
        s = self.simulationState
        mask = s["ghostIndices"] == -1
        cl = (lambda t: torch.clone(t)) if copy else (lambda t: t)
        p = {"UID": s["UID"][mask]}
        for k in ("fluidPosition", "fluidVelocity", "fluidDensity", "fluidSupport",
                  "fluidRestDensity", "fluidArea", "fluidPressure", "fluidAcceleration"):
            p[k] = cl(s[k][mask])
        if self.config["simulation"]["densityScheme"] == "continuum":
            p["dpdt"] = cl(s["dpdt"][mask])
        p["numParticles"] = p["fluidPosition"].shape[0]
        p["realParticles"] = p["numParticles"]
        for k in ("dt", "time", "timestep"):
            p[k] = s[k]
        for module in self.modules:
            module.saveState(p, copy=copy)
        return p

    def setupSimulationState(self, perennialState):
        """simulationBase.py:535-561."""
        s = {"UID": torch.clone(perennialState["UID"])}
        for k in ("fluidPosition", "fluidVelocity", "fluidDensity", "fluidSupport",
                  "fluidRestDensity", "fluidArea", "fluidPressure", "fluidAcceleration"):
            s[k] = torch.clone(perennialState[k])
        if self.config["simulation"]["densityScheme"] == "continuum":
            s["dpdt"] = torch.clone(perennialState["dpdt"])
        s["ghostIndices"] = torch.full_like(s["UID"], -1)
        for k in ("numParticles", "realParticles", "dt", "time", "timestep"):
            s[k] = perennialState[k]
        for module in self.modules:
            module.setupSimulationState(perennialState)
        return s

    def integrateValues(self, dt, dudt, dxdt, dpdt):
        """simulationBase.py:564-569."""
        self.simulationState["fluidVelocity"] = self.perennialState["fluidVelocity"] + dt * dudt
        self.simulationState["fluidPosition"] = self.perennialState["fluidPosition"] + dt * dxdt
        if self.config["simulation"]["densityScheme"] == "continuum":
            self.simulationState["fluidDensity"] = (self.perennialState["fluidDensity"] +
                                                    dt * dpdt / self.config["fluid"]["restDensity"])
        self.simulationState["fluidAcceleration"] = dudt

    def shiftParticles(self):
        """Step 14 of integrate (simulationBase.py:708-732): search and
        densities at the integrated positions, then the shifting module's
        shift.  The surface mask and normals are those of the last
        right-hand side, which the state still holds."""
        s = self.simulationState
        if not hasattr(self, "shiftingModule"):
            raise AttributeError("shifting is enabled but this scheme has no shifting module")
        self.neighborSearch.search(s, self)
        self.boundaryModule.boundaryFilterNeighborhoods(s, self)
        self.boundaryModule.boundaryNeighborhoodSearch(s, self)
        self.sphDensity.evaluate(s, self)
        self.boundaryModule.evalBoundaryDensity(s, self)
        self.shiftingModule.shift(s, self)
        s["fluidPosition"] += s["fluidUpdate"]

    def integrate(self):
        """One time step (simulationBase.py:632-782)."""
        self.perennialState = self.saveState()
        self.resetState()
        self.simulationState = self.setupSimulationState(self.perennialState)
        continuum = self.config["simulation"]["densityScheme"] == "continuum"
        dt = self.simulationState["dt"]
        scheme = self.config["integration"]["scheme"]
        if scheme == "explicitEuler":
            dudt, dxdt, dpdt = self.timestep()
            self.integrateValues(dt, dudt, dxdt, dpdt)
        elif scheme == "semiImplicitEuler":
            dudt, dxdt, dpdt = self.timestep()
            self.integrateValues(dt, dudt, dxdt + dt * dudt, dpdt)
        elif scheme in ("PECE", "PEC"):
            if scheme == "PECE":
                dudt, dxdt, dpdt = self.timestep()
            else:
                p = self.perennialState
                dudt, dxdt, dpdt = p["fluidAcceleration"], p["fluidVelocity"], p.get("dpdt")
            self.integrateValues(dt, dudt, dxdt, dpdt)
            dudt2, dxdt2, dpdt2 = self.timestep()
            self.integrateValues(dt, 0.5 * (dudt + dudt2), 0.5 * (dxdt + dxdt2),
                                 0.5 * (dpdt + dpdt2) if continuum else None)
        elif scheme == "RK4":
            k1 = self.timestep()
            self.simulationState = self.setupSimulationState(self.perennialState)
            self.integrateValues(0.5 * dt, *k1)
            k2 = self.timestep()
            self.simulationState = self.setupSimulationState(self.perennialState)
            self.integrateValues(0.5 * dt, *k2)
            k3 = self.timestep()
            self.simulationState = self.setupSimulationState(self.perennialState)
            self.integrateValues(dt, *k3)
            k4 = self.timestep()
            comb = [k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i] if k1[i] is not None else None
                    for i in range(3)]
            self.integrateValues(1 / 6 * dt, *comb)
        else:
            raise ValueError(f"unknown integration scheme {scheme}")
        if self.config["shifting"]["enabled"]:
            self.shiftParticles()
        self.perennialState["time"] += self.simulationState["dt"]
        self.perennialState["timestep"] += 1
        self.simulationState["time"] = self.perennialState["time"]
        self.simulationState["timestep"] = self.perennialState["timestep"]
        if getattr(self, "exporter", None) is not None:
            self.exporter.write_step(self)
