"""The oracle's neighbour search against the reference's own compiled CPU
search (cppSrc/neighSearch.cpp:172-331, built into oracle/_ref by
oracle/build_ref.py) through the golden fixtures it wrote, and live where
the compiled reference is present.  The near-tie case fixes the pair
predicate: pairs at distance exactly h whose fp32 squared distance rounds
below h^2 are not pairs in the reference (fl32(sqrt(s)) < h), although a
double-precision sqrt would accept them."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref_driver as R

GOLD = Path(__file__).resolve().parent / "golden" / "search_ref.npz"
CASES = ["uniform", "clustered", "variable", "neartie"]


def oracle_self(x, h):
    hmax = np.float32(h.max())
    qx, qy, cx, _ = O.grid_params(np.zeros(2, np.float32), np.ones(2, np.float32), hmax)
    return O.search(x, h, x, (qx, qy), hmax, cx, 1, self_mode=True).pairs()


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference_fixture(gold, case):
    x, h = gold[f"{case}_x"], gold[f"{case}_h"]
    r, c = oracle_self(x, h)
    assert np.array_equal(r, gold[f"{case}_rows"]) and np.array_equal(c, gold[f"{case}_cols"])


def test_near_tie_predicate_is_fp32_sqrt(gold):
    x, h = gold["neartie_x"], gold["neartie_h"]
    pairs = set(zip(gold["neartie_rows"].tolist(), gold["neartie_cols"].tolist()))
    hs = h[0]
    assert float(np.float32(hs * hs)) < float(hs) ** 2
    # points 0/1 and 2/3 sit exactly h apart with exact coordinate differences
    for a, b in ((0, 1), (2, 3)):
        d = x[a] - x[b]
        s = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1]))
        assert np.sqrt(np.float64(s)) < np.float64(hs)      # a double sqrt would pair them
        assert (a, b) not in pairs and (b, a) not in pairs


def test_oracle_foreign_matches_reference_fixture(gold):
    x, h = gold["uniform_x"], gold["uniform_h"]
    xq, hq = gold["foreign_xq"], gold["foreign_hq"]
    hmax = np.float32(h.max())
    qx, qy, cx, _ = O.grid_params(np.zeros(2, np.float32), np.ones(2, np.float32), hmax)
    r, c = O.search(xq, hq, x, (qx, qy), hmax, cx, 1).pairs()
    assert np.array_equal(r, gold["foreign_rows"]) and np.array_equal(c, gold["foreign_cols"])


@pytest.mark.skipif(not R.available(), reason="reference not compiled (oracle/build_ref.py)")
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(100 + seed)
    x = rng.uniform(0, 1, (4000, 2)).astype(np.float32)
    x[:50] = x[50:100]                                   # duplicates
    h = (rng.uniform(0.5, 1.0, len(x)) * np.float32(0.035)).astype(np.float32)
    r, c = R.self_search(x, h, (0.0, 0.0), (1.0, 1.0))
    ro, co = oracle_self(x, h)
    assert np.array_equal(r, ro) and np.array_equal(c, co)
