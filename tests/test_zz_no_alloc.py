"""tsm_pairs_noalloc (the entry point that allocates no device memory)
against tsm_pairs on the same inputs, through the C ABI: two calls in a row,
the second with a smaller input, for every measure; sets with long
trajectories (column strips, a reduced resident grid); a call too large for
the fixed scratch is refused with TSM_E_ARG."""
import ctypes

import numpy as np
import pytest

from tests.workloads import EPS, lognormal_set, uniform_pairs

pytestmark = pytest.mark.gpu

MEASURES = ["dtw", "edr", "lcss", "erp", "frechet", "hausdorff", "stedr", "stlcss", "cdds", "sar"]


def _decl(L):
    return L.tsm_pairs_noalloc


def _call(fn, m, d, pa, pb, prm, counted, bufs=None):
    import torch

    from trajsimimeasures_amd import _lib, engine
    n = pa.numel()
    if bufs is None:
        bufs = (torch.empty(n, dtype=torch.float64, device="cuda"),
                torch.empty(n, dtype=torch.int32, device="cuda") if counted else None)
    out, cnt = bufs
    out.fill_(-7.0)
    if cnt is not None:
        cnt.fill_(-7)
    s = d.cstruct()
    _lib.check(fn(_lib.measure_id(m), ctypes.byref(s), pa.data_ptr(), pb.data_ptr(), n, ctypes.byref(prm),
                  out.data_ptr(), cnt.data_ptr() if cnt is not None else None, engine._stream_ptr()))
    torch.cuda.synchronize()
    return out.cpu().numpy(), (cnt.cpu().numpy() if cnt is not None else None)


def _same(a, b):
    assert np.array_equal(a[0], b[0], equal_nan=True)
    if a[1] is not None:
        assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("measure", MEASURES)
@pytest.mark.parametrize("direct", [0, -1])
def test_noalloc_equals_tsm_pairs(gpu, measure, direct):
    """Two calls in a row, the second smaller.  With direct DP launches off
    (direct = 0) both DP calls plan through the fixed scratch; by default
    the smaller one is a direct launch, which needs none of it."""
    import torch

    from trajsimimeasures_amd import _lib, engine
    from trajsimimeasures_amd.data import with_timestamps
    L = _lib.lib()
    noalloc = _decl(L)
    print(f"[case] noalloc {measure} direct={direct}", flush=True)
    _lib.check(L.tsm_tune_set_direct_pairs(direct))
    try:
        _noalloc_two_calls(L, noalloc, measure, torch, engine, with_timestamps)
    finally:
        _lib.check(L.tsm_tune_set_direct_pairs(-1))


def _noalloc_two_calls(L, noalloc, measure, torch, engine, with_timestamps):
    from trajsimimeasures_amd import _lib
    ts, pa, pb = uniform_pairs(8000, seed=61)  # more pairs than a direct launch takes
    ts = with_timestamps(np.random.default_rng(62), ts)
    d = engine.DeviceTrajSet(ts)
    if measure == "sar":
        prm = _lib.params(EPS, time_step=10, target_length=10, replicas=64)
    else:
        prm = _lib.params(EPS, gpoint=(0.5, 0.5), delta=60.0)
    counted = _lib.measure_id(measure) in _lib.COUNTED
    pa_d = torch.as_tensor(pa, dtype=torch.int32, device="cuda")
    pb_d = torch.as_tensor(pb, dtype=torch.int32, device="cuda")
    small = slice(0, 1001)
    ref_big = _call(L.tsm_pairs, measure, d, pa_d, pb_d, prm, counted)
    ref_small = _call(L.tsm_pairs, measure, d, pb_d[small], pa_d[small], prm, counted)
    nb, ns = pa.size, small.stop
    b_big = (torch.empty(nb, dtype=torch.float64, device="cuda"),
             torch.empty(nb, dtype=torch.int32, device="cuda") if counted else None)
    b_small = (torch.empty(ns, dtype=torch.float64, device="cuda"),
               torch.empty(ns, dtype=torch.int32, device="cuda") if counted else None)
    # the runtime places the fixed scratch array in device memory when the
    # library first asks for its address: one warm-up call before measuring
    _call(noalloc, measure, d, pa_d[:10], pb_d[:10], prm, counted)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    got_big = _call(noalloc, measure, d, pa_d, pb_d, prm, counted, b_big)
    got_small = _call(noalloc, measure, d, pb_d[small], pa_d[small], prm, counted, b_small)
    free1 = torch.cuda.mem_get_info()[0]
    _same(got_big, ref_big)
    _same(got_small, ref_small)
    assert free1 >= free0, "the no-allocation entry point took device memory"


def test_noalloc_long_trajectories(gpu):
    """Lengths up to 1,000 points: column strips whose boundary vectors share
    the fixed scratch with the plan (fewer resident waves)."""
    import torch

    from trajsimimeasures_amd import _lib, engine
    L = _lib.lib()
    noalloc = _decl(L)
    print("[case] noalloc long trajectories", flush=True)
    ts = lognormal_set(6000, seed=63, lo=10, hi=1000)
    rng = np.random.default_rng(64)
    pa = rng.integers(0, ts.n, 30000)
    pb = rng.integers(0, ts.n, 30000)
    long_ = np.argsort(-ts.lengths)[:40]
    pa[:40], pb[:40] = long_, long_[::-1]
    d = engine.DeviceTrajSet(ts)
    pa_d = torch.as_tensor(pa, dtype=torch.int32, device="cuda")
    pb_d = torch.as_tensor(pb, dtype=torch.int32, device="cuda")
    for m in ("dtw", "edr"):
        prm = _lib.params(EPS)
        counted = m == "edr"
        for sl in (slice(0, 30000), slice(0, 700)):
            ref = _call(L.tsm_pairs, m, d, pa_d[sl], pb_d[sl], prm, counted)
            got = _call(noalloc, m, d, pa_d[sl], pb_d[sl], prm, counted)
            _same(got, ref)


def test_noalloc_refuses_oversized_calls(gpu):
    import torch

    from trajsimimeasures_amd import _lib, engine
    L = _lib.lib()
    noalloc = _decl(L)
    print("[case] noalloc oversized", flush=True)
    ts, pa, pb = uniform_pairs(100, seed=65)
    d = engine.DeviceTrajSet(ts)
    n = 9_000_000  # 36 MB of task slots: beyond the 32 MiB scratch
    idx = torch.zeros(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    s = d.cstruct()
    prm = _lib.params(EPS)
    rc = noalloc(0, ctypes.byref(s), idx.data_ptr(), idx.data_ptr(), n, ctypes.byref(prm), out.data_ptr(), None,
                 engine._stream_ptr())
    assert rc == -1 and b"no-allocation" in L.tsm_last_error()
    # the next call still works
    pa_d = torch.as_tensor(pa, dtype=torch.int32, device="cuda")
    pb_d = torch.as_tensor(pb, dtype=torch.int32, device="cuda")
    _same(_call(noalloc, "dtw", d, pa_d, pb_d, prm, False), _call(L.tsm_pairs, "dtw", d, pa_d, pb_d, prm, False))
