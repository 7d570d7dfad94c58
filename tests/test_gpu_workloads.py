"""Full-size runs of the query x database and all-pairs workloads, checked
through a random sample of pairs against the FP64 oracle and through
size-independent properties; plus the FP64 drift rerun on long-tail pairs.

  query x database: 1,000 queries x 100,000 trajectories (1e8 pairs),
    lengths lognormal (median 60, sigma 0.8) clipped to [10, 1000]; DTW, EDR
  all-pairs: upper triangle over 10,000 trajectories (49,995,000 pairs),
    lengths 98% uniform 10..500 and 2% uniform 1000..2000; DTW, Hausdorff;
    FP64 GPU rerun of the DTW and ERP pairs that touch a trajectory over
    1,000 points
Both full-size workloads are also captured into a hipGraph (grid and
triangle addressing, planner included) whose replays must equal the eager
call bit for bit.
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


def _replays_equal_eager(g, eager, eager_counts=None, reps=2):
    """Replays of graph g equal the eager result bit for bit, with eager
    work between the replays; returns the best replay time in seconds."""
    import torch
    best = float("inf")
    for _ in range(reps):
        g.out.fill_(-1.0)
        if g.counts is not None:
            g.counts.fill_(-1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        g.launch()
        g.stream.synchronize()
        best = min(best, time.perf_counter() - t)
        assert torch.equal(g.out, eager.reshape(-1))
        if eager_counts is not None:
            assert torch.equal(g.counts, eager_counts.reshape(-1))
        x = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        del x
    return best


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
        # the same grid captured into a hipGraph and replayed
        cnt = engine.grid(m, dset, 0, nq, nq, nd, eps=EPS, return_counts=True)[1] if m == "edr" else None
        g = engine.GridGraph(m, dset, 0, nq, nq, nd, eps=EPS)
        dt_g = _replays_equal_eager(g, out, cnt)
        print(f"[workload] query x db {m} graph replay: {nq * nd / dt_g:.3e} pairs/s", flush=True)
        g.close()
        del out, g, cnt


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
        g = engine.TriuGraph(m, dset, n)
        dt_g = _replays_equal_eager(g, out)
        print(f"[workload] all-pairs {m} graph replay: {iu0.size / dt_g:.3e} pairs/s", flush=True)
        g.close()
        del out, g
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
    # the same for ERP, whose boundary is a sum of gap distances (core/erp.py:44-49)
    g = ts.xy.mean(0)
    e32, e64, edrift = drivers.fp64_drift("erp", dset, pa, pb, gpoint=g)
    print(f"[workload] FP64 rerun of {pa.size} long-tail ERP pairs: max relative FP32 drift {edrift:.3e}", flush=True)
    ref, _ = O.pairs("erp", ts.xy, ts.offsets, pa[k], pb[k], gpoint=g)
    assert np.max(np.abs(e64[k] - ref) / ref) <= 1e-12
    assert np.max(np.abs(e32[k] - ref) / ref) <= 1e-4
    assert edrift <= 1e-4
