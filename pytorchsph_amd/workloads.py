"""The benchmark workloads as concrete synthetic inputs (fixed seeds).

  C1  delta-SPH dam break: tank [0,4]x[0,2], block 1.0 x 0.5 at dx = 0.01
      (5,000 fluid particles), 3 solid boundary layers with mirrored ghosts,
      jitter U(+-0.01 dx) from default_rng(0), RK4, MOG diffusion.
  C2  neighbour search alone: 4,000,000 points in [0,1)^2, half uniform,
      half 8 Gaussian clusters (sigma 0.05), support for 20 mean neighbours
      (default_rng(1)).
  C3  DFSPH dam break: block 1.0 x 1.0 at dx = 0.001 (1,000,000 particles),
      Akinci boundary, tau = 1e-3, default_rng(2).
  C4  batched: 256 scenes, counts log-uniform in [1e3, 2e5] from
      default_rng(3), dam break / colliding drops and delta-SPH / DFSPH
      alternating.
  C5  serving: Poisson arrivals (lambda = 20 / s, default_rng(4)), sizes
      log-uniform 2k-500k, 100-2,000 steps each, schemes 50/50.
"""
from __future__ import annotations

import math

import numpy as np

from . import scenes as S


def c1_scene():
    return S.dam_break(dx=0.01, block=(1.0, 0.5), scheme="dsph", seed=0)


def c2_points(n=4_000_000):
    x = S.point_cloud(n, seed=1)
    h = S.support_for_mean_neighbors(x, 20.0)
    return x, h


def c3_scene(dx=0.001):
    return S.dam_break(dx=dx, block=(1.0, 1.0), scheme="dfsph", seed=2)


def c4_specs(n_scenes=256, lo=1e3, hi=2e5, seed=3):
    """(kind, scheme, n_target, seed) of every batched scene."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(lo), math.log(hi), n_scenes)).astype(int)
    specs = []
    for k, n in enumerate(counts):
        kind = "dambreak" if k % 2 == 0 else "drops"
        scheme = "dsph" if (k // 2) % 2 == 0 else "dfsph"
        specs.append((kind, scheme, int(n), seed * 1000 + k))
    return specs


def c4_scenes(n_scenes=256, lo=1e3, hi=2e5, seed=3):
    return [S.scene_for_count(kind, scheme, n, sd)
            for kind, scheme, n, sd in c4_specs(n_scenes, lo, hi, seed)]


def c5_jobs(window_s=60.0, lam=20.0, seed=4, lo=2e3, hi=5e5, steps_lo=100, steps_hi=2000):
    """Poisson arrival times with job sizes / step counts / schemes."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / lam)
        if t > window_s:
            break
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        steps = int(math.exp(rng.uniform(math.log(steps_lo), math.log(steps_hi))))
        scheme = "dsph" if rng.uniform() < 0.5 else "dfsph"
        kind = "dambreak" if rng.uniform() < 0.5 else "drops"
        jobs.append(dict(t=t, n=n, steps=steps, scheme=scheme, kind=kind,
                         seed=int(rng.integers(1 << 30))))
    return jobs
