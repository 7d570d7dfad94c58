"""CPU-side checks: the C-ABI library loads and exports every function
include/tsm.h declares, argument errors come back as error codes without a
device, and the data layer (packing, generator) behaves."""
import ctypes
import os
import re

import numpy as np
import pytest

from trajsimimeasures_amd import _lib
from trajsimimeasures_amd.data import (TrajSet, lengths_lognormal, lengths_uniform, lengths_uniform_with_tail,
                                       pack_list, pack_padded, random_walks)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(name="tsm.h"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tsm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert L.tsm_abi_version() == 4
    nn = header_functions("tsm_nn.h")
    for n in nn:
        assert hasattr(L, n), n
    assert set(nn) == set(_lib.NN_EXPORTS)
    cl = header_functions("tsm_cluster.h")
    for n in cl:
        assert hasattr(L, n), n
    assert set(cl) == set(_lib.CLUSTER_EXPORTS)


def test_params_default():
    p = _lib.ParamsC()
    _lib.lib().tsm_params_default(ctypes.byref(p))
    assert (p.eps, p.eps_mode, p.gx, p.gy, p.lanes, p.precision) == (0.0, 0, 0.0, 0.0, 0, 0)
    # SAR: config.py:56-57 and core/sar.py:248
    assert (p.delta, p.time_step, p.target_length, p.replicas) == (0.0, 10, 10, 64)
    assert ctypes.sizeof(_lib.ParamsC) == 64


def test_argument_errors_without_device():
    L = _lib.lib()
    s = _lib.TrajSetC(0, 0, 0, 5, 5, 1, 0.0)
    prm = _lib.params()
    rc = L.tsm_pairs_triu(99, ctypes.byref(s), 3, ctypes.byref(prm), None, None, None)
    assert rc == -1 and b"unknown measure" in L.tsm_last_error()
    rc = L.tsm_pairs_triu(0, ctypes.byref(s), 6, ctypes.byref(prm), None, None, None)
    assert rc == -1 and b"triangle" in L.tsm_last_error()
    rc = L.tsm_pairs(0, None, None, None, 1, ctypes.byref(prm), None, None, None)
    assert rc == -1
    rc = L.tsm_pairs(0, ctypes.byref(s), None, None, 1, ctypes.byref(prm), None, None, None)
    assert rc == -1 and b"null pair" in L.tsm_last_error()
    rc = L.tsm_pairs_grid(0, ctypes.byref(s), 0, 5, 0, 5, ctypes.byref(prm), None, None, None)
    assert rc == -1
    xy = np.zeros((3, 2))
    offs = np.array([0, 2, 4], dtype=np.int64)  # does not end at npts
    pa = np.zeros(1, np.int32)
    out = np.zeros(1)
    rc = L.tsm_pairs_host(0, xy.ctypes.data, None, offs.ctypes.data, 2, 3, pa.ctypes.data, pa.ctypes.data, 1,
                          ctypes.byref(prm), out.ctypes.data, None)
    assert rc == -1 and b"offsets" in L.tsm_last_error()
    offs = np.array([0, 0, 3], dtype=np.int64)  # empty trajectory
    rc = L.tsm_pairs_host(0, xy.ctypes.data, None, offs.ctypes.data, 2, 3, pa.ctypes.data, pa.ctypes.data, 1,
                          ctypes.byref(prm), out.ctypes.data, None)
    assert rc == -1 and b"empty" in L.tsm_last_error()
    with pytest.raises(_lib.TsmError):
        _lib.check(-1)


def test_padded_host_argument_errors_without_device():
    """tsm_pairs_padded_host checks measure, point width, lengths and
    pointers before any device work (the batch pointers here are never
    dereferenced), and core.* turns these into ValueError."""
    L = _lib.lib()
    fake = ctypes.c_void_p(16)
    prm = _lib.params(0.1)
    lens = np.array([2, 3], dtype=np.int64)
    out = np.zeros(2)

    def call(measure=_lib.DTW, l1=lens, L1=3, l2=lens, L2=3, B=2, dim=2, t=fake, o=out):
        return L.tsm_pairs_padded_host(measure, t, l1.ctypes.data, L1, t, l2.ctypes.data, L2, B, dim,
                                       ctypes.byref(prm), o.ctypes.data if o is not None else None, None)

    assert call(B=0) == 0
    for kw, msg in ((dict(measure=99), b"unknown measure"), (dict(dim=4), b"2 (x, y) or 3"),
                    (dict(measure=_lib.STEDR), b"need [batch, seq, 3]"), (dict(L1=2), b"every length"),
                    (dict(l2=np.array([0, 1], dtype=np.int64)), b"every length"), (dict(L2=0), b"padded lengths"),
                    (dict(t=None), b"null"), (dict(o=None), b"null"), (dict(B=-1), b"negative")):
        assert call(**kw) == -1 and msg in L.tsm_last_error(), kw
    from trajsimimeasures_amd import core
    t = np.zeros((2, 3, 2))
    with pytest.raises(ValueError, match="every length"):
        core.dtw(t, [3, 4], t, [1, 1])
    with pytest.raises(ValueError, match="every length"):
        core.dtw(t, [3, 0], t, [1, 1])
    with pytest.raises(ValueError, match="lengths for a batch"):
        core.dtw(t, [3], t, [1, 1])
    with pytest.raises(ValueError):
        core.hausdorff(np.zeros((2, 3, 4)), [1, 1], np.zeros((2, 3, 4)), [1, 1])


def test_nonfinite_origin_rejected_without_device():
    """The FP32 conversion rejects a NaN or infinite origin before any
    device work (the point pointers are never dereferenced).  The library is
    built with -fno-honor-nans, under which a plain isfinite / bit-pattern
    test of a double loses its NaN half; this pins the test it uses."""
    L = _lib.lib()
    fake = ctypes.c_void_p(16)
    for ox, oy in ((float("nan"), 0.0), (0.0, float("nan")), (float("inf"), 0.0), (0.0, float("-inf")),
                   (-float("nan"), 1.0)):
        rc = L.tsm_xy_to_fp32_centred(fake, fake, 1, ox, oy, None)
        assert rc == -1 and b"origin must be finite" in L.tsm_last_error(), (ox, oy)


def test_packed_host_rejects_nonfinite_without_device():
    """tsm_pairs_host refuses non-finite coordinates and timestamps before any
    device work, as the padded entry point does: the kernels assume finite
    inputs."""
    L = _lib.lib()
    prm = _lib.params()
    offs = np.array([0, 2, 4], dtype=np.int64)
    pa = np.zeros(1, np.int32)
    pb = np.ones(1, np.int32)
    out = np.zeros(1)
    for bad in (float("nan"), float("inf"), -float("inf"), -float("nan")):
        xy = np.zeros((4, 2))
        xy[3, 1] = bad
        rc = L.tsm_pairs_host(0, xy.ctypes.data, None, offs.ctypes.data, 2, 4, pa.ctypes.data, pb.ctypes.data, 1,
                              ctypes.byref(prm), out.ctypes.data, None)
        assert rc == -1 and b"coordinates must be finite" in L.tsm_last_error(), bad
        t = np.arange(4.0)
        t[1] = bad
        rc = L.tsm_pairs_host(6, np.zeros((4, 2)).ctypes.data, t.ctypes.data, offs.ctypes.data, 2, 4,
                              pa.ctypes.data, pb.ctypes.data, 1, ctypes.byref(prm), out.ctypes.data, None)
        assert rc == -1 and b"timestamps must be finite" in L.tsm_last_error(), bad


def test_sar_argument_errors_without_device():
    """SAR needs timestamps and positive time_step / target_length /
    replicas; the checks run before any device work (the pointers here are
    never dereferenced)."""
    L = _lib.lib()
    fake = ctypes.c_void_p(16)
    s = _lib.TrajSetC(fake, fake, fake, 5, 5, 1, 0.0, None)
    prm = _lib.params(0.1)
    rc = L.tsm_pairs(_lib.SAR, ctypes.byref(s), fake, fake, 1, ctypes.byref(prm), fake, None, None)
    assert rc == -1 and b"timestamps" in L.tsm_last_error()
    s.t64 = fake
    for field in ("time_step", "target_length", "replicas"):
        bad = _lib.params(0.1)
        setattr(bad, field, 0)
        rc = L.tsm_pairs(_lib.SAR, ctypes.byref(s), fake, fake, 1, ctypes.byref(bad), fake, None, None)
        assert rc == -1 and b"SAR" in L.tsm_last_error(), field
    with pytest.raises(ValueError):
        _lib.params(0.1, time_step=2.5)
    with pytest.raises(ValueError):
        _lib.params(0.1, target_length=0)
    assert _lib.params(0.1, time_step=30.0, target_length=4, replicas=1).time_step == 30


def test_kmedoids_argument_errors_without_device():
    """Argument checks of the k-medoids entry point run before any device
    work (the matrix pointer is never dereferenced)."""
    L = _lib.lib()
    fake = ctypes.c_void_p(16)
    for n, k, med in ((10, 0, [0]), (10, 65, list(range(65))), (3, 4, [0, 1, 2, 3]), (10, 2, [3, 3]),
                      (10, 2, [1, 10])):
        m = np.array(med, dtype=np.int64)
        rc = L.tsm_kmedoids_fasterpam(fake, n, k, m.ctypes.data, 100, None, None, None, None, None)
        assert rc == -1 and b"kmedoids" in L.tsm_last_error(), (n, k, med)
    from trajsimimeasures_amd import cluster
    assert list(cluster.initial_medoids(10, 3, "first")) == [0, 1, 2]
    assert list(cluster.initial_medoids(100, 4, "random", 5)) == list(np.random.RandomState(5).choice(100, 4, False))
    with pytest.raises(ValueError):
        cluster.initial_medoids(10, [1, 1])
    with pytest.raises(ValueError):
        cluster.initial_medoids(10, 3, "build")


def test_params_validation():
    with pytest.raises(ValueError):
        _lib.params(eps_mode="cube")
    with pytest.raises(ValueError):
        _lib.params(precision="fp16")
    with pytest.raises(ValueError):
        _lib.params(gpoint=[1, 2, 3])
    with pytest.raises(ValueError):
        _lib.measure_id("RSTS")
    assert _lib.measure_id("dfrechet") == _lib.FRECHET


def test_pack_padded_roundtrip():
    rng = np.random.default_rng(0)
    lens = np.array([3, 1, 5])
    t = np.zeros((3, 5, 2))
    for i, n in enumerate(lens):
        t[i, :n] = rng.random((n, 2))
    ts = pack_padded(t, lens)
    assert ts.n == 3 and list(ts.lengths) == [3, 1, 5]
    for i, n in enumerate(lens):
        assert np.array_equal(ts.traj(i), t[i, :n])
    with pytest.raises(ValueError):
        pack_padded(t, [3, 1, 6])
    with pytest.raises(ValueError):
        pack_padded(t, [3, 0, 5])
    sub = ts.subset([2, 0])
    assert np.array_equal(sub.traj(0), ts.traj(2)) and np.array_equal(sub.traj(1), ts.traj(0))
    cat = TrajSet.concat([ts, sub])
    assert cat.n == 5 and np.array_equal(cat.traj(4), ts.traj(0))
    pl = pack_list([ts.traj(1), ts.traj(2)])
    assert list(pl.lengths) == [1, 5]


def test_random_walk_generator():
    rng = np.random.default_rng(0)
    lens = lengths_uniform(rng, 500, 20, 200)
    assert lens.min() >= 20 and lens.max() <= 200
    ts = random_walks(rng, lens)
    assert ts.n == 500 and np.array_equal(ts.lengths, lens)
    assert ts.xy.min() >= 0.0 and ts.xy.max() <= 1.0
    # reproducible from the seed, and equal to a point-by-point walk over the
    # same draws (starts, headings, turns, steps in that order)
    r1 = np.random.default_rng(7)
    lens7 = lengths_uniform(r1, 50, 1, 40)
    a = random_walks(r1, lens7)
    r2 = np.random.default_rng(7)
    lengths_uniform(r2, 50, 1, 40)
    p0 = r2.random((50, 2))
    th0 = r2.random(50) * (2.0 * np.pi)
    turn = r2.normal(0.0, 0.3, size=int(lens7.sum()))
    step = np.exp(r2.normal(np.log(0.002) - 0.125, 0.5, size=int(lens7.sum())))
    o = 0
    for t in range(50):
        x, y, cs = p0[t, 0], p0[t, 1], 0.0
        assert a.xy[o, 0] == x and a.xy[o, 1] == y
        for k in range(1, int(lens7[t])):
            cs += turn[o + k]
            th = th0[t] + cs
            x = min(max(x + step[o + k] * np.cos(th), 0.0), 1.0)
            y = min(max(y + step[o + k] * np.sin(th), 0.0), 1.0)
            assert a.xy[o + k, 0] == x and a.xy[o + k, 1] == y
        o += int(lens7[t])
    # step lengths: lognormal with mean 0.002 (clipping only shortens steps)
    steps = np.linalg.norm(np.diff(ts.xy, axis=0), axis=1)
    inner = np.ones(steps.size, bool)
    inner[ts.offsets[1:-1] - 1] = False
    assert 0.0017 < steps[inner].mean() < 0.0021


def test_length_samplers():
    rng = np.random.default_rng(0)
    ln = lengths_lognormal(rng, 20000)
    assert ln.min() >= 10 and ln.max() <= 1000 and 55 <= np.median(ln) <= 65
    lt = lengths_uniform_with_tail(rng, 20000)
    tail = lt >= 1000
    assert 0.015 < tail.mean() < 0.025 and lt[~tail].max() <= 500 and lt.max() <= 2000


def test_dropin_rejects_shape_mismatch():
    from trajsimimeasures_amd import core
    with pytest.raises(AssertionError):
        core.dtw(np.zeros((2, 3, 2)), [1, 1], np.zeros((2, 4, 2)), [1, 1])
    with pytest.raises(AssertionError):
        core.edr(np.zeros((2, 3, 2)), [1, 1], np.zeros((3, 3, 2)), [1, 1, 1], 0.1)


def test_pad_lists_to_array():
    from trajsimimeasures_amd.core import pad_lists_to_array
    arr = pad_lists_to_array([[[1, 2]], [[3, 4], [5, 6]]])
    assert arr.shape == (2, 2, 2) and arr.dtype == np.float64
    assert np.array_equal(arr[0], [[1, 2], [0, 0]])


def test_timestamps_and_st_packing():
    from trajsimimeasures_amd.data import with_timestamps
    rng = np.random.default_rng(3)
    ts = with_timestamps(rng, random_walks(rng, lengths_uniform(rng, 30, 1, 20)))
    for i in range(ts.n):
        t = ts.t[ts.offsets[i]:ts.offsets[i + 1]]
        assert 0 <= t[0] < 3600 and np.all(np.diff(t) >= 10.0 - 1e-9) and np.all(np.diff(t) <= 20.0 + 1e-9)
    sub = ts.subset([4, 2])
    assert np.array_equal(sub.t[:sub.lengths[0]], ts.t[ts.offsets[4]:ts.offsets[5]])
    cat = TrajSet.concat([ts, sub])
    assert cat.t.size == cat.xy.shape[0]
    # [B, L, 3] padded batches and [n, 3] lists carry the timestamps
    lens = np.array([2, 3])
    arr = np.zeros((2, 3, 3))
    arr[0, :2] = [[0.1, 0.2, 5.0], [0.3, 0.4, 6.0]]
    arr[1] = [[0.5, 0.6, 1.0], [0.7, 0.8, 2.0], [0.9, 1.0, 3.0]]
    p = pack_padded(arr, lens)
    assert np.array_equal(p.t, [5.0, 6.0, 1.0, 2.0, 3.0]) and p.xy.shape == (5, 2)
    pl = pack_list([arr[0, :2], arr[1]])
    assert np.array_equal(pl.t, p.t) and np.array_equal(pl.xy, p.xy)
    with pytest.raises(ValueError):
        TrajSet(p.xy, p.offsets, p.t[:-1])


def test_st_dropin_rejects_2d_points():
    from trajsimimeasures_amd import core
    with pytest.raises(ValueError):
        core.stedr(np.zeros((2, 3, 2)), [1, 1], np.zeros((2, 3, 2)), [1, 1], 0.1, 60)


def test_autotune_cost_model_mirror():
    """The tuner's copy of the library's lanes cost model: the fewest lanes
    whose register columns (24 per lane, 12 with two lanes) hold the pair."""
    from trajsimimeasures_amd import autotune as A
    got = [A.model_lanes(24, nc) for nc in range(1, 257)]
    want = [2 if nc <= 24 else 4 if nc <= 96 else 8 if nc <= 192 else 16 for nc in range(1, 257)]
    assert got == want
    assert [A.model_lanes(16, nc) for nc in (16, 17, 64, 65, 128, 129)] == [2, 4, 4, 8, 8, 16]
    assert all(A.fits(24, A.model_lanes(24, nc), nc) for nc in range(1, 385))


def test_dropin_lanes_hint():
    """threads_per_traj of the reference's drop-in calls: 2, 4, 8, 16 are
    lanes hints, the reference default 64 (and anything else) leaves the
    launch shape to the library."""
    from trajsimimeasures_amd.core._batch import lanes_hint
    assert [lanes_hint(t) for t in (2, 4, 8, 16)] == [2, 4, 8, 16]
    assert [lanes_hint(t) for t in (64, 32, 1, 0, None, "x")] == [0] * 6
