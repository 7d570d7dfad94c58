"""The reference's pair datasets over the package's own HDF5 reader.

``DatasetST_h5`` restates utilities/data_loader.py:52-114: open a dataset
file (preprocessing/pkl2h5py.py:16-27 layout), keep the trajectories with at
least ``traj_len_min`` points, give each one longer than ``traj_len_max`` a
random length in [traj_len_min, traj_len_max] (it is truncated to its first
that many points), then draw ``num_trajs`` first members and ``num_trajs``
second members with replacement.  The draws use Python's Mersenne Twister
seeded with ``seed`` and make the same calls in the same order as the
reference (one ``randint`` per long trajectory in file order, then two
``choices`` of k = num_trajs), so the same file and seed give the same pairs.
The reference seeds the interpreter-wide ``random`` module; this class uses
its own ``random.Random(seed)``, which yields the same sequence without
touching the global state.

``DatasetSynthetic`` restates data_loader.py:117-148: uniform points in
[0,1]^2, lengths uniform in [traj_len_min, traj_len_max], drawn per item
from numpy's global generator, with cumulative integer timestamps for the
spatio-temporal form.

Items are the reference's (traj1, traj2) pairs of Python lists
(``__getitem__``); ``packed()`` gives the same pairs as one packed
``TrajSet`` plus index arrays, which is what the engine and the drivers take
without building any list.

Coordinates stay float64 as stored.  Mercator coordinates are ~1e7 m: the
engine centres them before its FP32 copy is made (``DeviceTrajSet``), so no
precision is lost to their magnitude.
"""
from __future__ import annotations

import math
import random

import numpy as np

from .data import TrajSet
from .h5 import H5File

__all__ = ["DatasetST_h5", "DatasetSynthetic", "read_h5_set", "sample_h5_indices"]


def sample_h5_indices(trajs_len, num_trajs: int, traj_len_min: int, traj_len_max: int, seed):
    """The (index, length) pairs DatasetST_h5 draws (data_loader.py:67-73,
    105-111): a list of ((i1, n1), (i2, n2)) of length ``num_trajs``."""
    rng = random.Random(seed)
    trajs_len = [int(v) for v in trajs_len]
    kept = [(i, n) for i, n in enumerate(trajs_len) if n >= traj_len_min]
    kept = [(i, rng.randint(traj_len_min, traj_len_max)) if n > traj_len_max else (i, n) for i, n in kept]
    if not len(kept) >= math.sqrt(num_trajs):
        raise AssertionError(f"{len(kept)} trajectories of at least {traj_len_min} points; "
                             f"{num_trajs} pairs need at least sqrt({num_trajs})")
    first = rng.choices(kept, k=num_trajs)
    second = rng.choices(kept, k=num_trajs)
    return list(zip(first, second))


def _gather(f: H5File, items, use_temporal: bool) -> TrajSet:
    """Packed set of the given (index, length) items, in that order."""
    lens = np.fromiter((n for _, n in items), dtype=np.int64, count=len(items))
    offsets = np.zeros(len(items) + 1, np.int64)
    np.cumsum(lens, out=offsets[1:])
    xy = np.empty((int(offsets[-1]), 2), np.float64)
    t = np.empty(int(offsets[-1]), np.float64) if use_temporal else None
    for k, (i, n) in enumerate(items):
        a = f["/trajs_merc/%s" % i].read(n)
        if a.ndim != 2 or a.shape[1] < 2 or a.shape[0] < n:
            raise ValueError(f"/trajs_merc/{i}: shape {a.shape}, {n} points of (x, y) needed")
        xy[offsets[k]:offsets[k + 1]] = a[:n, :2]
        if use_temporal:
            ts = f["/trajs_ts/%s" % i].read(n)
            if ts.shape[0] < n:
                raise ValueError(f"/trajs_ts/{i}: {ts.shape[0]} timestamps, {n} needed")
            t[offsets[k]:offsets[k + 1]] = ts[:n]
    return TrajSet(xy, offsets, t)


def read_h5_set(file_path: str, use_temporal: bool = False, indices=None):
    """Trajectories of a dataset file as one packed TrajSet (all of them, or
    those in ``indices``, whole), with the file's ``merc_range`` and
    ``ts_range`` (None when absent)."""
    with H5File(file_path) as f:
        lens = f["/trajs_len"][:].astype(np.int64)
        idx = range(len(lens)) if indices is None else [int(i) for i in indices]
        ts = _gather(f, [(i, int(lens[i])) for i in idx], use_temporal)
        merc = np.asarray(f.attrs["merc_range"], np.float64)
        tsr = np.asarray(f.attrs["ts_range"]) if "ts_range" in f.attrs else None
    return ts, merc, tsr


class DatasetST_h5:
    """Spatial (or spatio-temporal) pairs from a dataset file
    (data_loader.py:52-114).  ``self.merc_range`` / ``self.ts_range`` are the
    file's attributes; items are (traj1, traj2) lists of [x, y] (or
    [x, y, t]) points."""

    def __init__(self, file_path: str, use_temporal: bool, num_trajs, traj_len_min, traj_len_max, seed):
        self.use_temporal = use_temporal
        self.traj_len_min = traj_len_min
        self.traj_len_max = traj_len_max
        self.data_length = num_trajs
        with H5File(file_path) as f:
            self.merc_range = np.asarray(f.attrs["merc_range"])
            self.ts_range = np.asarray(f.attrs["ts_range"]) if "ts_range" in f.attrs else None
            trajs_len = f["/trajs_len"][:].tolist()
            sampled = sample_h5_indices(trajs_len, num_trajs, traj_len_min, traj_len_max, seed)
            # every distinct (index, length) item is read once
            items, where = [], {}
            for iv in (iv for pair in sampled for iv in pair):
                if iv not in where:
                    where[iv] = len(items)
                    items.append(iv)
            self._set = _gather(f, items, use_temporal)
        self._pa = np.fromiter((where[p[0]] for p in sampled), dtype=np.int32, count=len(sampled))
        self._pb = np.fromiter((where[p[1]] for p in sampled), dtype=np.int32, count=len(sampled))
        self.sampled_indices = sampled

    def packed(self):
        """(TrajSet of the distinct sampled items, first indices, second
        indices): item k is the pair (set[pa[k]], set[pb[k]])."""
        return self._set, self._pa, self._pb

    def _traj(self, j):
        s = self._set
        a, b = s.offsets[j], s.offsets[j + 1]
        if not self.use_temporal:
            return s.xy[a:b].tolist()
        return np.column_stack([s.xy[a:b], s.t[a:b]]).tolist()

    def __getitem__(self, index):
        return (self._traj(int(self._pa[index])), self._traj(int(self._pb[index])))

    def __len__(self):
        return self.data_length


class DatasetSynthetic:
    """Uniform random pairs (data_loader.py:117-148): each item draws fresh
    lengths and points from numpy's global generator; ``seed`` seeds Python's
    ``random`` module, as in the reference, which the items do not use."""

    def __init__(self, use_temporal: bool, num_trajs, traj_len_min, traj_len_max, seed):
        random.seed(seed)
        self.use_temporal = use_temporal
        self.traj_len_min = traj_len_min
        self.traj_len_max = traj_len_max
        self.num_trajs = num_trajs
        self.merc_range = [0.0, 1.0, 0.0, 1.0]

    def __getitem__(self, index):
        # the draws (and their order) define the pairs: length, points, length,
        # points, then the two sets of time steps of 1..59 s
        lens = []
        pts = []
        for _ in range(2):
            lens.append(np.random.randint(self.traj_len_min, self.traj_len_max + 1))
            pts.append(np.random.uniform(size=(lens[-1], 2)).astype(np.float64))
        if not self.use_temporal:
            return (pts[0].tolist(), pts[1].tolist())
        stamps = [np.cumsum(np.random.randint(1, 60, size=n)) for n in lens]
        return tuple(np.column_stack([p, t]).tolist() for p, t in zip(pts, stamps))

    def __len__(self):
        return self.num_trajs
