"""Physics modules with the reference's classes, verbs and simulationState
keys (src/modules/*.py); their per-pair sums run in the gfx950 library."""
