"""Shared body of the drop-in per-measure functions: hand the padded batch
to libtsm.so's padded host entry point, return float64 [B] like the
reference (core/dtw.py:56-77)."""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib

__all__ = ["run_padded", "lanes_hint"]


def lanes_hint(threads_per_traj) -> int:
    """The reference's block size per pair (config.py:45) is a hint here: 2,
    4, 8 and 16 select that many lanes per pair where they hold the pair's
    columns; anything else (the reference default 64 included) leaves the
    launch shape to the library (a direct launch for a small batch, else the
    planner with the tuned lanes table)."""
    try:
        t = int(threads_per_traj)
    except (TypeError, ValueError):
        return 0
    return t if t in (2, 4, 8, 16) else 0


def run_padded(measure: int, t1, len1, t2, len2, *, eps=0.0, eps_mode="sqrt", gpoint=None, threads_per_traj=64,
               precision="fp32", return_counts=False, delta=0.0, sar=None):
    """One batch through tsm_pairs_padded_host: the padded arrays go to the
    device as they are and are compacted there.  Argument errors (lengths
    outside [1, padded length], non-finite values inside the lengths, shapes)
    raise ValueError; the batch sizes must agree (AssertionError, as the
    reference's own assert)."""
    t1 = np.ascontiguousarray(t1, dtype=np.float64)
    t2 = np.ascontiguousarray(t2, dtype=np.float64)
    st = measure in _lib.TIMED
    width = 3 if st else 2
    if t1.ndim != 3 or t2.ndim != 3 or (st and (t1.shape[2] != 3 or t2.shape[2] != 3)):
        raise ValueError(f"t1 and t2 must be [batch, seq, {width}] arrays")
    if t1.shape[2] not in (2, 3) or t2.shape[2] not in (2, 3):
        raise ValueError("points must have 2 (x, y) or 3 (x, y, t) values")
    if t1.shape[2] != t2.shape[2]:  # (x, y) with (x, y, t): the measure reads (x, y) only
        t1 = np.ascontiguousarray(t1[..., :2])
        t2 = np.ascontiguousarray(t2[..., :2])
    assert t1.shape[0] == t2.shape[0]
    B = t1.shape[0]
    if B == 0:
        out = np.zeros(0, dtype=np.float64)
        return (out, np.zeros(0, dtype=np.int32)) if return_counts else out
    l1 = np.ascontiguousarray(np.asarray(len1).reshape(-1), dtype=np.int64)
    l2 = np.ascontiguousarray(np.asarray(len2).reshape(-1), dtype=np.int64)
    if l1.size != B or l2.size != B:
        raise ValueError(f"{l1.size} and {l2.size} lengths for a batch of {B} pairs")
    out = np.empty(B, dtype=np.float64)
    cnt = np.empty(B, dtype=np.int32) if return_counts else None
    if sar is not None:  # (time_eps, target_length, threads_per_traj) of core/sar.py:247-248
        prm = _lib.params(eps, time_step=sar[0], target_length=sar[1], replicas=sar[2])
    else:
        prm = _lib.params(eps, eps_mode, gpoint, lanes_hint(threads_per_traj), precision, delta)
    L = _lib.lib()
    rc = L.tsm_pairs_padded_host(measure, t1.ctypes.data, l1.ctypes.data, t1.shape[1], t2.ctypes.data,
                                 l2.ctypes.data, t2.shape[1], B, t1.shape[2], ctypes.byref(prm), out.ctypes.data,
                                 cnt.ctypes.data if cnt is not None else None)
    if rc == -1:  # TSM_E_ARG
        raise ValueError(L.tsm_last_error().decode())
    _lib.check(rc)
    return (out, cnt) if return_counts else out
