"""The padded-batch host entry point (tsm_pairs_padded_host, behind every
core.* drop-in) against the packed one (tsm_pairs_host) on the same pairs.

The padded arrays go to the device as they are and are compacted there; the
packed entry point receives the same points already compacted on the host.
Both take the FP32 origin from the same bounding box and run the same
kernels, so the bar is bit-identical values and counts for every measure.
The padding is filled with NaN: it must never be read.
"""
import ctypes

import numpy as np
import pytest

from tests.workloads import EPS, uniform_pairs

pytestmark = pytest.mark.gpu

MEASURES = ["dtw", "edr", "lcss", "erp", "frechet", "hausdorff", "stedr", "stlcss", "cdds", "sar"]
SAR = (10, 10, 64)  # time_eps, target_length, threads_per_traj


def _pairs(npairs, seed, lo=1, hi=120):
    from trajsimimeasures_amd.data import TrajSet, with_timestamps
    ts, pa, pb = uniform_pairs(npairs, lo=lo, hi=hi, seed=seed)
    ts = with_timestamps(np.random.default_rng(seed + 1), ts)
    return TrajSet.concat([ts.subset(pa), ts.subset(pb)])


def _padded(ts, first, B, width, dim, fill=np.nan):
    arr = np.full((B, width, dim), fill)
    lens = ts.lengths[first:first + B].copy()
    for k in range(B):
        s, e = ts.offsets[first + k], ts.offsets[first + k + 1]
        arr[k, :e - s, :2] = ts.xy[s:e]
        if dim == 3:
            arr[k, :e - s, 2] = ts.t[s:e]
    return arr, lens


def _params(m):
    from trajsimimeasures_amd import _lib
    if m == "sar":
        return _lib.params(EPS, time_step=SAR[0], target_length=SAR[1], replicas=SAR[2])
    return _lib.params(EPS, delta=60.0)


def _packed_host(m, ts, B):
    """tsm_pairs_host on the compacted set: pair k is (k, B + k)."""
    from trajsimimeasures_amd import _lib
    mid = _lib.measure_id(m)
    timed = mid in _lib.TIMED
    xy = np.ascontiguousarray(ts.xy)
    offs = np.ascontiguousarray(ts.offsets, dtype=np.int64)
    t = np.ascontiguousarray(ts.t) if timed else None
    pa = np.arange(B, dtype=np.int32)
    pb = pa + np.int32(B)
    out = np.empty(B)
    cnt = np.empty(B, dtype=np.int32)
    prm = _params(m)
    _lib.check(_lib.lib().tsm_pairs_host(mid, xy.ctypes.data, t.ctypes.data if timed else None, offs.ctypes.data,
                                         2 * B, xy.shape[0], pa.ctypes.data, pb.ctypes.data, B, ctypes.byref(prm),
                                         out.ctypes.data, cnt.ctypes.data))
    return out, cnt


def _padded_host(m, t1, l1, t2, l2):
    from trajsimimeasures_amd import _lib
    from trajsimimeasures_amd.core._batch import run_padded
    mid = _lib.measure_id(m)
    if m == "sar":
        return run_padded(mid, t1, l1, t2, l2, eps=EPS, sar=SAR, return_counts=True)
    return run_padded(mid, t1, l1, t2, l2, eps=EPS, delta=60.0, return_counts=True)


@pytest.mark.parametrize("measure", MEASURES)
def test_padded_equals_packed(gpu, measure):
    """B = 6,000 pairs (12,000 trajectories: more than one pass of the
    compaction's grid-stride loop), lengths 1..120, the two batches padded
    to different widths, NaN in every padding slot."""
    print(f"[case] padded host vs packed host measure={measure}", flush=True)
    B = 6000
    ts = _pairs(B, seed=11)
    dim = 3  # timestamps present; read only by the timed measures
    t1, l1 = _padded(ts, 0, B, int(ts.lengths[:B].max()) + 3, dim)
    t2, l2 = _padded(ts, B, B, int(ts.lengths[B:].max()) + 17, dim)
    got, gcnt = _padded_host(measure, t1, l1, t2, l2)
    ref, rcnt = _packed_host(measure, ts, B)
    assert np.array_equal(got, ref, equal_nan=True), f"{measure}: {int((got != ref).sum())} values differ"
    if measure in ("edr", "lcss", "stedr", "stlcss"):
        assert np.array_equal(gcnt, rcnt)


def test_padded_two_column_points_and_the_dropins(gpu):
    """[B, L, 2] batches (no timestamps) through core.dtw / core.edr equal
    the packed entry point; a [B, L, 3] batch with NaN timestamps gives the
    same spatial results (the timestamps are not read)."""
    from trajsimimeasures_amd import core
    B = 700
    ts = _pairs(B, seed=3, lo=1, hi=60)
    width = int(ts.lengths.max())
    t1, l1 = _padded(ts, 0, B, width, 2, fill=0.0)
    t2, l2 = _padded(ts, B, B, width, 2, fill=0.0)
    ref, _ = _packed_host("dtw", ts, B)
    assert np.array_equal(core.dtw(t1, l1, t2, l2), ref)
    ref_e, _ = _packed_host("edr", ts, B)
    assert np.array_equal(core.edr(t1, l1, t2, l2, EPS), ref_e)
    u1, _ = _padded(ts, 0, B, width, 3)
    u2, _ = _padded(ts, B, B, width, 3)
    u1[:, :, 2] = np.nan
    u2[:, :, 2] = np.nan
    assert np.array_equal(core.dtw(u1, l1, u2, l2), ref)


def test_padded_rejects_nonfinite_inside_lengths(gpu):
    from trajsimimeasures_amd import core
    B = 50
    ts = _pairs(B, seed=5, lo=2, hi=30)
    width = int(ts.lengths.max())
    t1, l1 = _padded(ts, 0, B, width, 3)
    t2, l2 = _padded(ts, B, B, width, 3)
    bad = t1.copy()
    bad[7, l1[7] - 1, 0] = np.inf
    with pytest.raises(ValueError, match="finite"):
        core.dtw(bad, l1, t2, l2)
    bad = t2.copy()
    bad[3, 0, 2] = np.nan
    with pytest.raises(ValueError, match="finite"):
        core.stedr(t1, l1, bad, l2, EPS, 60.0)
    # the same NaN timestamp is not read by a spatial measure
    ref, _ = _packed_host("dtw", ts, B)
    assert np.array_equal(core.dtw(t1, l1, bad, l2), ref)
    # and the call after a rejected one is unaffected
    assert np.array_equal(core.dtw(t1, l1, t2, l2), ref)
