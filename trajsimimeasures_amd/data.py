"""Trajectory sets: the packed layout, padded <-> packed conversion, and the
seeded random-walk generator used for every synthetic workload.

Layout (host and HBM alike): a trajectory set is one contiguous float64 array
``xy`` of shape [N_points, 2] (x, y interleaved) plus ``offsets`` int64
[n_traj + 1]; trajectory t is ``xy[offsets[t]:offsets[t+1]]``.  The engine
keeps an FP32 copy of ``xy`` next to the FP64 one on the device.

The reference passes zero-padded batches [B, L, 2] plus length vectors
(core/dtw.py:56-62, utilities/tool_funcs.py:230-234); ``pack_padded`` turns
those into this layout without touching padding.

Random walk (one PCG64 stream per workload, ``numpy.random.default_rng``):
start uniform in [0,1]^2, initial heading uniform in [0, 2*pi), heading turns
by N(0, 0.3 rad) per step, step length lognormal with mean 0.002 and log-sigma
0.5 (log-mean ln(0.002) - 0.5**2/2), every position clipped to [0,1]^2.  Draw
order is fixed (lengths, starts, headings, turns, step lengths) so a workload
is reproducible from its seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = [
    "TrajSet",
    "with_timestamps",
    "pack_padded",
    "pack_list",
    "random_walks",
    "lengths_uniform",
    "lengths_lognormal",
    "lengths_uniform_with_tail",
]


@dataclass
class TrajSet:
    """Packed trajectories: ``xy`` float64 [N, 2], ``offsets`` int64 [n + 1],
    and optionally per-point timestamps ``t`` float64 [N] (the third column
    of the reference's spatio-temporal [batch, seq, 3] inputs)."""

    xy: np.ndarray
    offsets: np.ndarray
    t: np.ndarray | None = None

    def __post_init__(self):
        self.xy = np.ascontiguousarray(self.xy, dtype=np.float64).reshape(-1, 2)
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.int64)
        if self.offsets.ndim != 1 or self.offsets.size < 1:
            raise ValueError("offsets must be a 1-D array of n_traj + 1 entries")
        if self.offsets[0] != 0 or self.offsets[-1] != self.xy.shape[0]:
            raise ValueError("offsets must start at 0 and end at the number of points")
        if np.any(np.diff(self.offsets) < 0):
            raise ValueError("offsets must be non-decreasing")
        if self.t is not None:
            self.t = np.ascontiguousarray(self.t, dtype=np.float64).reshape(-1)
            if self.t.size != self.xy.shape[0]:
                raise ValueError("t must hold one timestamp per point")

    @property
    def n(self) -> int:
        return self.offsets.size - 1

    @property
    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)

    def traj(self, i: int) -> np.ndarray:
        return self.xy[self.offsets[i]:self.offsets[i + 1]]

    def subset(self, idx) -> "TrajSet":
        idx = np.asarray(idx, dtype=np.int64)
        lens = self.lengths[idx]
        offs = np.zeros(idx.size + 1, dtype=np.int64)
        np.cumsum(lens, out=offs[1:])
        if idx.size == 0:
            return TrajSet(np.zeros((0, 2)), offs, None if self.t is None else np.zeros(0))
        # gather point indices of every selected trajectory
        starts = self.offsets[idx]
        rel = np.arange(offs[-1], dtype=np.int64) - np.repeat(offs[:-1], lens)
        pts = np.repeat(starts, lens) + rel
        return TrajSet(self.xy[pts], offs, None if self.t is None else self.t[pts])

    @staticmethod
    def concat(sets) -> "TrajSet":
        sets = list(sets)
        xy = np.concatenate([s.xy for s in sets], axis=0) if sets else np.zeros((0, 2))
        lens = np.concatenate([s.lengths for s in sets]) if sets else np.zeros(0, np.int64)
        offs = np.zeros(lens.size + 1, dtype=np.int64)
        np.cumsum(lens, out=offs[1:])
        t = None
        if sets and all(s.t is not None for s in sets):
            t = np.concatenate([s.t for s in sets])
        return TrajSet(xy, offs, t)


def pack_padded(t: np.ndarray, lens) -> TrajSet:
    """Pack a zero-padded batch [B, L, 2] with lengths [B] (reference layout,
    core/dtw.py:58-59) into a TrajSet; padding rows are dropped.  A
    [B, L, 3] batch (x, y, timestamp; core/stedr.py:62-64) also fills t."""
    t = np.asarray(t)
    if t.ndim != 3 or t.shape[2] not in (2, 3):
        raise ValueError(f"expected a [B, L, 2] or [B, L, 3] array, got shape {t.shape}")
    lens = np.asarray(lens, dtype=np.int64).reshape(-1)
    if lens.size != t.shape[0]:
        raise ValueError(f"{lens.size} lengths for a batch of {t.shape[0]} trajectories")
    if lens.size and (lens.min() < 1 or lens.max() > t.shape[1]):
        raise ValueError("every length must be in [1, padded length]")
    mask = np.arange(t.shape[1])[None, :] < lens[:, None]
    offs = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    pts = np.asarray(t, dtype=np.float64)[mask]
    return TrajSet(pts[:, :2], offs, pts[:, 2] if t.shape[2] == 3 else None)


def pack_list(trajs) -> TrajSet:
    """Pack a list of [n_i, 2] (or, all of them, [n_i, 3] with timestamps)
    arrays / lists."""
    arrs = [np.asarray(a, dtype=np.float64) for a in trajs]
    width = 3 if arrs and all(a.ndim == 2 and a.shape[1] == 3 for a in arrs) else 2
    arrs = [a.reshape(-1, width) for a in arrs]
    lens = np.array([a.shape[0] for a in arrs], dtype=np.int64)
    offs = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    pts = np.concatenate(arrs, axis=0) if arrs else np.zeros((0, width))
    return TrajSet(pts[:, :2], offs, pts[:, 2] if width == 3 else None)


def lengths_uniform(rng: np.random.Generator, n: int, lo: int, hi: int) -> np.ndarray:
    """n lengths uniform on the integers lo..hi inclusive."""
    return rng.integers(lo, hi + 1, size=n).astype(np.int64)


def lengths_lognormal(rng: np.random.Generator, n: int, median: float = 60.0, sigma: float = 0.8,
                      lo: int = 10, hi: int = 1000) -> np.ndarray:
    """clip(rint(exp(N(ln median, sigma))), lo, hi)."""
    x = np.rint(np.exp(rng.normal(np.log(median), sigma, size=n)))
    return np.clip(x, lo, hi).astype(np.int64)


def lengths_uniform_with_tail(rng: np.random.Generator, n: int, lo: int = 10, hi: int = 500,
                              tail_frac: float = 0.02, tail_lo: int = 1000, tail_hi: int = 2000) -> np.ndarray:
    """(1 - tail_frac) of lengths uniform on lo..hi, the rest on tail_lo..tail_hi."""
    base = rng.integers(lo, hi + 1, size=n)
    tail = rng.integers(tail_lo, tail_hi + 1, size=n)
    is_tail = rng.random(n) < tail_frac
    return np.where(is_tail, tail, base).astype(np.int64)


STEP_MEAN = 0.002
STEP_SIGMA = 0.5
TURN_SIGMA = 0.3


def random_walks(rng: np.random.Generator, lengths) -> TrajSet:
    """Seeded random-walk trajectories with the given lengths (see module doc).

    All random draws happen up front in a fixed order (start points, initial
    headings, one turn and one step per point), then the walks advance one
    point index at a time over the trajectories still that long (length-
    sorted, so each index touches a prefix), so the cost is linear in the
    number of points whatever the length spread."""
    lens = np.asarray(lengths, dtype=np.int64).reshape(-1)
    if lens.size and lens.min() < 1:
        raise ValueError("lengths must be >= 1")
    n = lens.size
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    total = int(offs[-1])
    p0 = rng.random((n, 2))
    th0 = rng.random(n) * (2.0 * np.pi)
    turn = rng.normal(0.0, TURN_SIGMA, size=total)
    mu = np.log(STEP_MEAN) - 0.5 * STEP_SIGMA ** 2
    step = np.exp(rng.normal(mu, STEP_SIGMA, size=total))
    xy = np.empty((total, 2), dtype=np.float64)
    if n == 0:
        return TrajSet(xy, offs)
    order = np.argsort(-lens, kind="stable")
    slen = lens[order]
    base = offs[:-1][order]
    px = p0[order, 0].copy()
    py = p0[order, 1].copy()
    th = th0[order]
    cs = np.zeros(n)                     # running sum of turns since point 0
    xy[base, 0] = px
    xy[base, 1] = py
    # number of trajectories longer than k, for every k
    active = np.searchsorted(-slen, -np.arange(int(slen[0])), side="left")
    for k in range(1, int(slen[0])):
        m = int(active[k])
        idx = base[:m] + k
        cs[:m] += turn[idx]
        theta = th[:m] + cs[:m]
        st = step[idx]
        px[:m] = np.clip(px[:m] + st * np.cos(theta), 0.0, 1.0)
        py[:m] = np.clip(py[:m] + st * np.sin(theta), 0.0, 1.0)
        xy[idx, 0] = px[:m]
        xy[idx, 1] = py[:m]
    return TrajSet(xy, offs)


def with_timestamps(rng: np.random.Generator, ts: TrajSet, interval: float = 15.0, jitter: float = 5.0,
                    span: float = 3600.0) -> TrajSet:
    """The same trajectories with per-point timestamps in seconds: a start
    uniform on [0, span), then sampling intervals uniform on
    interval +- jitter (a GPS log's shape), drawn in point order."""
    n = ts.n
    start = rng.random(n) * span
    gaps = interval + jitter * (2.0 * rng.random(ts.xy.shape[0]) - 1.0)
    lens = ts.lengths
    first = ts.offsets[:-1]
    gaps[first[lens > 0]] = 0.0
    csum = np.cumsum(gaps)
    base = np.repeat(csum[first] if n else np.zeros(0), lens)
    t = np.repeat(start, lens) + (csum - base)
    return TrajSet(ts.xy, ts.offsets, t)
