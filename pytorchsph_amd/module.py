"""Module base classes, same hooks as the reference (src/module.py:4-56)."""


class Module:
    def getParameters(self):
        return None

    def initialize(self, config, state):
        return

    def finalize(self):
        return

    def resetState(self, simulationState):
        return

    def saveState(self, perennialState, copy):
        return

    def exportState(self, simulationState, simulation, group, mask):
        return

    def setupSimulationState(self, perennialState):
        return

    def __init__(self, identifier, moduleDescription):
        self.description = moduleDescription
        self.name = identifier
        self.identifier = identifier


class BoundaryModule(Module):
    """Boundary hooks used by the delta-SPH and DFSPH solvers; unimplemented
    hooks raise, as in the reference."""

    def __init__(self, identifier, moduleDescription):
        super().__init__(identifier, moduleDescription)

    def _missing(self, op):
        raise Exception(f"Operation {op} not implemented for {self.identifier}")

    def dfsphPrepareSolver(self, simulationState, simulation, density):
        self._missing("dfsphPrepareSolver")

    def dfsphBoundaryAccelTerm(self, simulationState, simulation, density):
        self._missing("dfsphBoundaryAccelTerm")

    def dfsphBoundaryPressureSum(self, simulationState, simulation, density):
        self._missing("dfsphBoundaryPressureSum")

    def dfsphBoundaryAlphaTerm(self, simulationState, simulation, density):
        self._missing("dfsphBoundaryAlphaTerm")

    def dfsphBoundarySourceTerm(self, simulationState, simulation, density):
        self._missing("dfsphBoundarySourceTerm")

    def evalBoundaryPressure(self, simulationState, simulation, density):
        self._missing("boundaryPressure")

    def evalBoundaryDensity(self, simulationState, simulation):
        self._missing("boundaryDensity")

    def evalBoundaryFriction(self, simulationState, simulation):
        self._missing("boundaryFriction")

    def boundaryNeighborhoodSearch(self, simulationState, simulation):
        self._missing("boundaryNeighborsearch")

    def boundaryFilterNeighborhoods(self, simulationState, simulation):
        return
