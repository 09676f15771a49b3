"""Writes tests/golden/search_ref.npz: small neighbour-search inputs and the
pair lists that the reference's own compiled CPU search returns for them
(oracle/_ref, built by oracle/build_ref.py from cppSrc/neighSearch.cpp).

Cases (all fp32, domain [0,1]^2, searchRadius 1):
  uniform    3,000 uniform points, h for ~20 neighbours
  clustered  8,000 points, 8 Gaussian clusters + uniform half
  variable   5,000 points, h_i in [0.3, 1] * hmax
  neartie    exact-difference pairs at distance h for supports h whose
             square rounds down in fp32 (fl32(sqrt(s)) == h: no pair,
             while sqrt in double would give one), points on cell faces,
             duplicates, a dense cell and the domain corners
  foreign    1,000 query points (own supports) against the uniform grid
Each case stores x, h (and xq, hq) and the lexicographically sorted
(rows, cols).

Usage: python tests/golden/make_search_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import ref_driver as R  # noqa: E402
from pytorchsph_amd import scenes as S  # noqa: E402


def near_tie_points():
    pts = []
    h = np.float32(0.05)
    found = []
    while len(found) < 3:
        if float(np.float32(h * h)) < float(h) * float(h):
            found.append(h)
        h = np.nextafter(h, np.float32(1))
    hsel = found[0]
    pts += [[0.0, 0.5], [hsel, 0.5], [0.5, 0.0], [0.5, hsel]]
    # +-3 ulp around h from an interior point (inexact differences)
    for k in range(-3, 4):
        d = hsel
        for _ in range(abs(k)):
            d = np.nextafter(d, np.float32(1 if k > 0 else 0))
        pts += [[0.3 + d, 0.7], [0.3, 0.7 - d]]
    pts += [[0.3, 0.7]]
    qmin = np.float32(0) - hsel
    faces = qmin + hsel * np.arange(3, 15, dtype=np.float32)
    pts += [[f, 0.21] for f in faces] + [[0.83, f] for f in faces]
    pts += [[0.61, 0.42]] * 4
    rng = np.random.default_rng(5)
    pts += list(0.9 + rng.uniform(0, 0.01, (40, 2)))
    pts += [[0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0]]
    x = np.array(pts, np.float32)
    return x, np.full(len(x), hsel, np.float32)


def main():
    out = {}
    rng = np.random.default_rng(21)
    x = rng.uniform(0, 1, (3000, 2)).astype(np.float32)
    h = np.full(len(x), np.float32(np.sqrt(20 / (np.pi * len(x)))), np.float32)
    cases = {"uniform": (x, h)}
    xc = S.point_cloud(8_000, seed=22)
    cases["clustered"] = (xc, np.full(len(xc), S.support_for_mean_neighbors(xc, 20.0, sample=4000),
                                      np.float32))
    xv = rng.uniform(0, 1, (5000, 2)).astype(np.float32)
    hv = (rng.uniform(0.3, 1.0, len(xv)) * np.float32(0.03)).astype(np.float32)
    hv[0] = np.float32(0.03)
    cases["variable"] = (xv, hv)
    cases["neartie"] = near_tie_points()
    for name, (px, ph) in cases.items():
        r, c = R.self_search(px, ph, (0.0, 0.0), (1.0, 1.0))
        out[f"{name}_x"], out[f"{name}_h"] = px, ph
        out[f"{name}_rows"], out[f"{name}_cols"] = r.astype(np.int32), c.astype(np.int32)
    xq = rng.uniform(-0.02, 1.02, (1000, 2)).astype(np.float32)
    hq = np.full(len(xq), h[0], np.float32)
    r, c = R.foreign_search(xq, hq, x, h, (0.0, 0.0), (1.0, 1.0))
    out["foreign_xq"], out["foreign_hq"] = xq, hq
    out["foreign_rows"], out["foreign_cols"] = r.astype(np.int32), c.astype(np.int32)
    path = Path(__file__).resolve().parent / "search_ref.npz"
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({path.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
