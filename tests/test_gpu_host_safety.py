"""Host entry points under concurrent callers, in-place refills of a set
behind a captured graph, and pairs longer than the set's stated max_len."""
import ctypes
import threading

import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import EPS, fixed_lengths, uniform_pairs

pytestmark = pytest.mark.gpu


def _host_call(ts, pa, pb, measure):
    from trajsimimeasures_amd import _lib
    L = _lib.lib()
    prm = _lib.params(EPS)
    pa32 = np.ascontiguousarray(pa, np.int32)
    pb32 = np.ascontiguousarray(pb, np.int32)
    out = np.empty(pa32.size)
    cnt = np.empty(pa32.size, np.int32)
    _lib.check(L.tsm_pairs_host(_lib.measure_id(measure), ts.xy.ctypes.data, None, ts.offsets.ctypes.data, ts.n,
                                int(ts.offsets[-1]), pa32.ctypes.data, pb32.ctypes.data, pa32.size,
                                ctypes.byref(prm), out.ctypes.data, cnt.ctypes.data))
    return out, cnt


def test_host_entry_points_from_two_threads(gpu):
    """Two host threads call tsm_pairs_host and the padded drop-in
    (core.dtw) on different inputs at once, several times; every result
    equals the same call made alone (ctypes releases the GIL, so the calls
    really overlap)."""
    from trajsimimeasures_amd import core
    print("[case] two host threads, tsm_pairs_host + core.dtw", flush=True)
    ts1, pa1, pb1 = uniform_pairs(3000, seed=41)
    ts2, pa2, pb2 = uniform_pairs(5000, lo=5, hi=300, seed=42)
    serial1 = _host_call(ts1, pa1, pb1, "edr")
    w = int(ts2.lengths.max())

    def padded(idx):
        arr = np.zeros((idx.size, w, 2))
        for k, i in enumerate(idx):
            arr[k, :ts2.lengths[i]] = ts2.traj(i)
        return arr, ts2.lengths[idx]

    t1, l1 = padded(pa2)
    t2, l2 = padded(pb2)
    serial2 = core.dtw(t1, l1, t2, l2)
    errors = []

    def worker(fn, want, reps=6):
        try:
            for _ in range(reps):
                got = fn()
                if isinstance(want, tuple):
                    ok = all(np.array_equal(g, s) for g, s in zip(got, want))
                else:
                    ok = np.array_equal(got, want)
                if not ok:
                    errors.append(fn)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(lambda: _host_call(ts1, pa1, pb1, "edr"), serial1)),
          threading.Thread(target=worker, args=(lambda: core.dtw(t1, l1, t2, l2), serial2))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    ref, _ = O.pairs("dtw", ts2.xy, ts2.offsets, pa2, pb2)
    assert np.max(np.abs(serial2 - ref) / np.maximum(ref, 1e-300)) <= 1e-5


@pytest.mark.parametrize("measure", ["edr", "lcss", "dtw", "erp"])
def test_refill_behind_a_graph(gpu, measure):
    """Coordinates moved in place (DeviceTrajSet.refill: FP64 copy, FP32 copy
    at the same origin, within the set's bound) are what the next replay
    computes on, and equal the oracle on the moved points; points beyond the
    bound are refused."""
    import torch

    from trajsimimeasures_amd import engine
    print(f"[case] refill behind graph {measure}", flush=True)
    ts, pa, pb = uniform_pairs(3000, seed=51)
    d = engine.DeviceTrajSet(ts)
    g = engine.PairsGraph(measure, d, pa, pb, eps=EPS, gpoint=(0.4, 0.6))
    # move every point towards the origin and jitter it: same lengths, new values
    rng = np.random.default_rng(5)
    o = np.asarray(d.origin)
    moved = o + 0.8 * (ts.xy - o) + rng.normal(0.0, 1e-3, ts.xy.shape)
    moved = np.clip(moved, o - d.coord_max, o + d.coord_max)
    d.refill(moved)
    torch.cuda.synchronize()
    g.launch()
    g.stream.synchronize()
    got = g.out.cpu().numpy()
    ref, _ = O.pairs(measure, moved, ts.offsets, pa, pb, eps=EPS, gpoint=(0.4, 0.6))
    if measure in ("edr", "lcss"):
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5
    with pytest.raises(ValueError):
        d.refill(ts.xy + 2.0 * d.coord_max)
    g.close()


def test_understated_max_len_gives_nan_and_minus_one(gpu):
    """A pair wider than one whole-wavefront strip (1,536 columns) whose set
    claims a smaller max_len is not computed: NaN, and count -1 for the
    counting measures."""
    import torch

    from trajsimimeasures_amd import _lib, engine
    print("[case] understated max_len", flush=True)
    ts, pa, pb = fixed_lengths([30, 2000, 40], [50, 1800, 60], seed=3)
    d = engine.DeviceTrajSet(ts)
    s = d.cstruct()
    s.max_len = 100  # understated: the middle pair needs strips
    n = pa.size
    pa_d = torch.as_tensor(pa, dtype=torch.int32, device="cuda")
    pb_d = torch.as_tensor(pb, dtype=torch.int32, device="cuda")
    ref = {m: O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=EPS)[0] for m in ("edr", "dtw")}
    for m in ("edr", "dtw"):
        out = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
        cnt = torch.full((n,), 7, dtype=torch.int32, device="cuda")
        prm = _lib.params(EPS)
        _lib.check(_lib.lib().tsm_pairs(_lib.measure_id(m), ctypes.byref(s), pa_d.data_ptr(), pb_d.data_ptr(), n,
                                        ctypes.byref(prm), out.data_ptr(), cnt.data_ptr() if m == "edr" else None,
                                        engine._stream_ptr()))
        got = out.cpu().numpy()
        assert np.isnan(got[1]), (m, got)
        keep = np.array([0, 2])
        if m == "edr":
            assert cnt.cpu().numpy()[1] == -1
            assert np.array_equal(got[keep], ref[m][keep])
        else:
            assert np.max(np.abs(got[keep] - ref[m][keep]) / ref[m][keep]) <= 1e-5
