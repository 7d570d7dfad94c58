"""Seeded synthetic workloads shared by the tests (random-walk generator of
trajsimimeasures_amd.data)."""
from __future__ import annotations

import numpy as np

from trajsimimeasures_amd.data import TrajSet, lengths_lognormal, lengths_uniform, random_walks

EPS = 0.0025  # EDR/LCSS epsilon of the 100k-pair six-measure workload


def uniform_pairs(npairs: int, lo: int = 20, hi: int = 200, seed: int = 0):
    """npairs independent pairs (2k, 2k+1), lengths uniform on lo..hi."""
    rng = np.random.default_rng(seed)
    ts = random_walks(rng, lengths_uniform(rng, 2 * npairs, lo, hi))
    pa = np.arange(0, 2 * npairs, 2, dtype=np.int64)
    return ts, pa, pa + 1


def lognormal_set(n: int, seed: int = 0, lo: int = 10, hi: int = 1000) -> TrajSet:
    rng = np.random.default_rng(seed)
    return random_walks(rng, lengths_lognormal(rng, n, lo=lo, hi=hi))


def fixed_lengths(lens_a, lens_b, seed: int = 0):
    """Pairs with prescribed lengths (pa[k] has lens_a[k] points)."""
    rng = np.random.default_rng(seed)
    la = np.asarray(lens_a, dtype=np.int64)
    lb = np.asarray(lens_b, dtype=np.int64)
    lens = np.empty(2 * la.size, dtype=np.int64)
    lens[0::2] = la
    lens[1::2] = lb
    ts = random_walks(rng, lens)
    pa = np.arange(0, lens.size, 2, dtype=np.int64)
    return ts, pa, pa + 1
