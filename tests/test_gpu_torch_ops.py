"""torch.ops.sph called with the reference's argument lists, as
src/modules/neighborSearch.py:114,138,160,170 calls its extension, returns
the pairs of the reference's own compiled search (tests/golden fixtures),
for device and host inputs and from TorchScript."""
import numpy as np
import pytest
import torch

from oracle import ref_driver
from pytorchsph_amd.torch_ops import neighborSearch
from tests.test_torch_ops import scripted_cuda_search, scripted_self_search

pytestmark = pytest.mark.gpu

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "search_ref.npz"
CASES = ["uniform", "clustered", "variable", "neartie"]


def structures(x, h, dev):
    """sortPositions + constructHashMap (neighborSearch.py:30-88), on `dev`."""
    s = ref_driver.hash_structures(x, h, (0, 0), (1, 1))
    t = {k: (v.to(dev) if torch.is_tensor(v) else v) for k, v in s.items()}
    t["x"] = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    t["h"] = torch.from_numpy(np.ascontiguousarray(h, np.float32)).to(dev)
    return t


def lexsorted(r, c):
    r, c = r.cpu().numpy().astype(np.int64), c.cpu().numpy().astype(np.int64)
    o = np.lexsort((c, r))
    return r[o], c[o]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("where", ["cuda", "cpu"])
def test_build_neighbor_list_cuda_matches_reference(cuda, case, where):
    g = np.load(GOLD)
    t = structures(g[f"{case}_x"], g[f"{case}_h"], where)
    i, j = neighborSearch.buildNeighborListCUDA(
        t["x"], t["h"], t["sx"], t["sh"], t["table"], t["cell_keys"], t["counts"], t["starts"],
        t["sort"], t["qmin"], np.float32(t["hmax"]).item(), t["cells_x"], t["P"], 1)
    assert i.device.type == where and i.dtype == torch.int32 and j.dtype == torch.int32
    r, c = i.cpu().numpy(), j.cpu().numpy()
    # already in the lexicographic order of the reference's finalize step
    assert np.all((np.diff(r) > 0) | ((np.diff(r) == 0) & (np.diff(c) > 0)))
    assert np.array_equal(r, g[f"{case}_rows"]) and np.array_equal(c, g[f"{case}_cols"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("where", ["cuda", "cpu"])
def test_build_neighbor_list_matches_reference(cuda, case, where):
    g = np.load(GOLD)
    t = structures(g[f"{case}_x"], g[f"{case}_h"], where)
    rows, cols = neighborSearch.buildNeighborList(
        t["sx"], t["sh"], t["table"], t["cell_keys"], t["counts"], t["starts"], t["sort"],
        t["cells_x"], t["P"], 1)
    assert rows.device.type == where
    r, c = lexsorted(rows, cols)
    assert np.array_equal(r, g[f"{case}_rows"]) and np.array_equal(c, g[f"{case}_cols"])


def test_build_neighbor_list_unsorted_per_particle_matches_reference(cuda):
    g = np.load(GOLD)
    t = structures(g["uniform_x"], g["uniform_h"], "cuda")
    xq = torch.from_numpy(g["foreign_xq"]).cuda()
    hq = torch.from_numpy(g["foreign_hq"]).cuda()
    rows, cols = neighborSearch.buildNeighborListUnsortedPerParticle(
        xq, hq, t["sx"], t["sh"], t["table"], t["cell_keys"], t["counts"], t["starts"],
        t["sort"], t["cells_x"], t["P"], t["qmin"], np.float32(t["hmax"]).item(), 1)
    r, c = lexsorted(rows, cols)
    assert np.array_equal(r, g["foreign_rows"]) and np.array_equal(c, g["foreign_cols"])


def test_scripted_callers_run_on_the_device(cuda):
    g = np.load(GOLD)
    t = structures(g["clustered_x"], g["clustered_h"], "cuda")
    rows, cols = scripted_self_search(t["sx"], t["sh"], t["table"], t["cell_keys"], t["counts"],
                                      t["starts"], t["sort"], t["cells_x"], t["P"])
    r, c = lexsorted(rows, cols)
    assert np.array_equal(r, g["clustered_rows"]) and np.array_equal(c, g["clustered_cols"])
    i, j = scripted_cuda_search(t["x"], t["h"], t["sx"], t["sh"], t["table"], t["cell_keys"],
                                t["counts"], t["starts"], t["sort"], t["qmin"],
                                np.float32(t["hmax"]).item(), t["cells_x"], t["P"])
    assert np.array_equal(i.cpu().numpy(), g["clustered_rows"])
    assert np.array_equal(j.cpu().numpy(), g["clustered_cols"])


def test_bad_arguments_raise(cuda):
    g = np.load(GOLD)
    t = structures(g["uniform_x"], g["uniform_h"], "cuda")
    with pytest.raises(RuntimeError, match="cumCell sums"):
        neighborSearch.buildNeighborList(t["sx"][:-1], t["sh"][:-1], t["table"], t["cell_keys"],
                                         t["counts"], t["starts"], t["sort"][:-1], t["cells_x"],
                                         t["P"], 1)
