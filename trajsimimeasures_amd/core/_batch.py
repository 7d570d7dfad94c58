"""Shared body of the drop-in per-measure functions: pack the padded batch,
call libtsm.so's host entry point, return float64 [B] like the reference
(core/dtw.py:56-77)."""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ..data import pack_padded

__all__ = ["run_padded", "lanes_hint"]


def lanes_hint(threads_per_traj) -> int:
    """The reference's block size per pair (config.py:45) is a hint here: 4
    and 8 select that many lanes per pair (16 maps to 8), anything else
    (the reference default 64 included) the autotuned table."""
    try:
        t = int(threads_per_traj)
    except (TypeError, ValueError):
        return 0
    return {4: 4, 8: 8, 16: 8}.get(t, 0)


def run_padded(measure: int, t1, len1, t2, len2, *, eps=0.0, eps_mode="sqrt", gpoint=None, threads_per_traj=64,
               precision="fp32", return_counts=False, delta=0.0, sar=None):
    t1 = np.asarray(t1, dtype=np.float64)
    t2 = np.asarray(t2, dtype=np.float64)
    st = measure in _lib.TIMED
    width = 3 if st else 2
    if t1.ndim != 3 or t2.ndim != 3 or (st and (t1.shape[2] != 3 or t2.shape[2] != 3)):
        raise ValueError(f"t1 and t2 must be [batch, seq, {width}] arrays")
    if not st:
        t1, t2 = t1[..., :2], t2[..., :2]
    assert t1.shape[0] == t2.shape[0]
    B = t1.shape[0]
    if B == 0:
        out = np.zeros(0, dtype=np.float64)
        return (out, np.zeros(0, dtype=np.int32)) if return_counts else out
    a = pack_padded(t1, len1)
    b = pack_padded(t2, len2)
    xy = np.ascontiguousarray(np.concatenate([a.xy, b.xy], axis=0))
    if not np.all(np.isfinite(xy)):
        raise ValueError("coordinates must be finite")
    offs = np.concatenate([a.offsets, b.offsets[1:] + a.offsets[-1]]).astype(np.int64)
    tt = np.ascontiguousarray(np.concatenate([a.t, b.t])) if st else None
    if st and not np.all(np.isfinite(tt)):
        raise ValueError("timestamps must be finite")
    pa = np.arange(B, dtype=np.int32)
    pb = pa + np.int32(B)
    out = np.empty(B, dtype=np.float64)
    cnt = np.empty(B, dtype=np.int32) if return_counts else None
    if sar is not None:  # (time_eps, target_length, threads_per_traj) of core/sar.py:247-248
        prm = _lib.params(eps, time_step=sar[0], target_length=sar[1], replicas=sar[2])
    else:
        prm = _lib.params(eps, eps_mode, gpoint, lanes_hint(threads_per_traj), precision, delta)
    L = _lib.lib()
    _lib.check(L.tsm_pairs_host(measure, xy.ctypes.data, tt.ctypes.data if st else None, offs.ctypes.data, 2 * B,
                                xy.shape[0], pa.ctypes.data,
                                pb.ctypes.data, B, ctypes.byref(prm), out.ctypes.data,
                                cnt.ctypes.data if cnt is not None else None))
    return (out, cnt) if return_counts else out
