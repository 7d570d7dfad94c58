"""Full-size runs of the query x database and all-pairs workloads, checked
through a random sample of pairs against the FP64 oracle and through
size-independent properties; plus the FP64 drift rerun on long-tail pairs.

  query x database: 1,000 queries x 100,000 trajectories (1e8 pairs),
    lengths lognormal (median 60, sigma 0.8) clipped to [10, 1000]; DTW, EDR
  all-pairs: upper triangle over 10,000 trajectories (49,995,000 pairs),
    lengths 98% uniform 10..500 and 2% uniform 1000..2000; DTW, Hausdorff;
    FP64 GPU rerun of the DTW pairs that touch a trajectory over 1,000 points
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import EPS

pytestmark = pytest.mark.gpu


def _sample_check(measure, ts, qa, qb, got, rng, n=1500, cnt=None):
    k = rng.choice(qa.size, size=min(n, qa.size), replace=False)
    ref, rc = O.pairs(measure, ts.xy, ts.offsets, qa[k], qb[k], eps=EPS)
    if measure in ("edr", "lcss"):
        assert np.array_equal(got[k], ref)
    else:
        rel = np.abs(got[k] - ref) / np.maximum(ref, 1e-300)
        assert rel.max() <= 1e-5, (measure, rel.max())


def test_query_by_database_1e8(gpu):
    import torch

    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    rng = np.random.default_rng(0)
    nq, nd = 1000, 100_000
    ts = random_walks(rng, lengths_lognormal(rng, nq + nd))
    dset = engine.DeviceTrajSet(ts)
    lens = ts.lengths.astype(np.float64)
    cells = lens[:nq].sum() * lens[nq:].sum()
    for m in ("dtw", "edr"):
        engine.grid(m, dset, 0, 10, nq, nd, eps=EPS)  # warm the workspace
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = engine.grid(m, dset, 0, nq, nq, nd, eps=EPS)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        print(f"[workload] query x db {m}: {nq * nd / dt:.3e} pairs/s, {cells / dt:.3e} cells/s", flush=True)
        v = out.reshape(-1).cpu().numpy()
        assert np.all(np.isfinite(v)) and v.min() >= 0
        qa = np.repeat(np.arange(nq), nd)
        qb = np.tile(np.arange(nq, nq + nd), nq)
        _sample_check(m, ts, qa, qb, v, rng)
        del out


def test_all_pairs_triangle_and_drift(gpu):
    import torch

    from trajsimimeasures_amd import drivers, engine
    from trajsimimeasures_amd.data import lengths_uniform_with_tail, random_walks
    rng = np.random.default_rng(0)
    n = 10_000
    ts = random_walks(rng, lengths_uniform_with_tail(rng, n))
    dset = engine.DeviceTrajSet(ts)
    iu0, iu1 = np.triu_indices(n, 1)
    lens = ts.lengths.astype(np.float64)
    for m in ("dtw", "hausdorff"):
        engine.triu(m, dset, 50)
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = engine.triu(m, dset, n)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        cells = float(np.sum(lens[iu0] * lens[iu1]))
        print(f"[workload] all-pairs {m}: {iu0.size / dt:.3e} pairs/s, {cells / dt:.3e} cells/s", flush=True)
        v = out.cpu().numpy()
        assert v.size == n * (n - 1) // 2 and np.all(np.isfinite(v)) and v.min() >= 0
        _sample_check(m, ts, iu0, iu1, v, rng)
        del out
    # FP64 rerun of the DTW pairs touching a trajectory longer than 1000 points
    long = ts.lengths > 1000
    sel = long[iu0] | long[iu1]
    pa, pb = iu0[sel].astype(np.int32), iu1[sel].astype(np.int32)
    v32, v64, drift = drivers.fp64_drift("dtw", dset, pa, pb)
    print(f"[workload] FP64 rerun of {pa.size} long-tail DTW pairs: max relative FP32 drift {drift:.3e}", flush=True)
    k = rng.choice(pa.size, 300, replace=False)
    ref, _ = O.pairs("dtw", ts.xy, ts.offsets, pa[k], pb[k])
    assert np.max(np.abs(v64[k] - ref) / ref) <= 1e-12  # the FP64 path matches the oracle
    assert drift <= 1e-4
