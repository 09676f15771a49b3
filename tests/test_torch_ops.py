"""The torch operator library (libsph_torch.so, TORCH_LIBRARY(sph)) without a
GPU: it loads, registers the reference's three search entry points with the
reference's argument lists (cppSrc/neighSearch.cpp:143-177,261-267), and
TorchScript compiles calls to them (the reference's pybind module cannot be
called from TorchScript)."""
import torch

from pytorchsph_amd.torch_ops import neighborSearch

REFERENCE_ARGS = {
    "buildNeighborListCUDA": ["queryParticles", "support", "sortedParticles", "sortedSupport",
                              "hashTable", "cellIndices", "cumCell", "cellSpan", "sort", "qMin",
                              "hMax", "cellsX", "hashMapLength", "searchRadius"],
    "buildNeighborList": ["queryParticles", "support", "hashTable", "cellIndices", "cumCell",
                          "cellSpan", "sort", "cellsX", "hashMapLength", "searchRadius"],
    "buildNeighborListUnsortedPerParticle": ["inputParticles", "inputSupport", "queryParticles",
                                             "support", "hashTable", "cellIndices", "cumCell",
                                             "cellSpan", "sort", "cellsX", "hashMapLength",
                                             "qMin", "hMax", "searchRadius"],
}


def test_ops_registered_with_reference_argument_lists():
    for name, args in REFERENCE_ARGS.items():
        schema = getattr(torch.ops.sph, name).default._schema
        assert [a.name for a in schema.arguments] == args, name
        assert str(schema.returns[0].type) == "List[Tensor]"
    assert neighborSearch.version() >= 1


@torch.jit.script
def scripted_self_search(x, h, table, ci, cc, cs, s, cells_x: int, P: int):
    return torch.ops.sph.buildNeighborList(x, h, table, ci, cc, cs, s, cells_x, P, 1)


@torch.jit.script
def scripted_cuda_search(xq, hq, x, h, table, ci, cc, cs, s, qmin, hmax: float, cells_x: int,
                         P: int):
    return torch.ops.sph.buildNeighborListCUDA(xq, hq, x, h, table, ci, cc, cs, s, qmin, hmax,
                                               cells_x, P, 1)


def test_torchscript_compiles_calls_to_the_ops():
    for fn in (scripted_self_search, scripted_cuda_search):
        assert "sph::" in fn.graph.str()
