"""Parity of SAR (core/sar.py) on the GPU against the oracle (the reference
GPU form, oracle/tsm_oracle.c tsmo_sar): equal values for increasing
timestamps (the data-parallel path), scrambled timestamps (the sequential
path), edge lengths and long stretches, every parameter regime (per-sample
counts with their threads_per_traj replication, bins fewer and more than a
wavefront), explicit / grid / triangle addressing and the padded drop-in."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import uniform_pairs

pytestmark = pytest.mark.gpu


def _st_set(n, lo, hi, seed, span=600.0):
    from trajsimimeasures_amd.data import with_timestamps
    ts, pa, pb = uniform_pairs(n, lo=lo, hi=hi, seed=seed)
    return with_timestamps(np.random.default_rng(seed + 1), ts, span=span), pa, pb


def _check(ts, pa, pb, got, deps, tstep, M, R):
    ref = O.pairs_sar(ts.xy, ts.t, ts.offsets, pa, pb, deps, tstep, M, R)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first {bad[:5]}: {got[bad[:5]]} vs {ref[bad[:5]]}"
    return ref


@pytest.mark.parametrize("tstep,M,R", [(10, 10, 64), (10, 1000, 64), (1, 10, 1), (30, 3, 64), (10, 100, 1),
                                       (7, 1, 64)])
def test_sar_random_pairs(gpu, tstep, M, R):
    from trajsimimeasures_amd import engine
    print(f"[case] sar random pairs time_eps={tstep} M={M} R={R}", flush=True)
    ts, pa, pb = _st_set(3000, 2, 150, seed=61)
    d = engine.DeviceTrajSet(ts)
    for deps in (0.003, 0.02, 0.2):
        got = engine.pairs("sar", d, pa, pb, eps=deps, time_step=tstep, target_length=M, replicas=R)
        ref = _check(ts, pa, pb, got.cpu().numpy(), deps, tstep, M, R)
    assert np.count_nonzero(ref) > 100  # the workload has overlaps and matches


def test_sar_scrambled_timestamps(gpu):
    """Decreasing timestamps take the sequential path; still equal."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import TrajSet
    print("[case] sar scrambled timestamps", flush=True)
    ts, pa, pb = _st_set(1500, 2, 100, seed=63)
    rng = np.random.default_rng(64)
    t = ts.t.copy()
    for i in rng.choice(ts.n, ts.n // 2, replace=False):  # swap two timestamps in half the trajectories
        o, n = ts.offsets[i], ts.offsets[i + 1] - ts.offsets[i]
        a, b = o + rng.integers(0, n, 2)
        t[a], t[b] = t[b], t[a]
    ts2 = TrajSet(ts.xy, ts.offsets, t)
    d = engine.DeviceTrajSet(ts2)
    for M, R in ((10, 64), (2, 1), (500, 64)):
        got = engine.pairs("sar", d, pa, pb, eps=0.02, time_step=10, target_length=M, replicas=R)
        _check(ts2, pa, pb, got.cpu().numpy(), 0.02, 10, M, R)


def test_sar_edges_long_and_identity(gpu):
    """Lengths 1 and 2, long trajectories (stretches of thousands of samples,
    bins spanning several 64-sample chunks), identical pairs, repeated
    timestamps, trajectories without any common time."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import lengths_uniform, random_walks, with_timestamps
    print("[case] sar edges", flush=True)
    rng = np.random.default_rng(65)
    lens = np.concatenate([[1, 2, 1, 2, 3000, 2500, 129, 65, 64], lengths_uniform(rng, 40, 1, 400)])
    ts = with_timestamps(rng, random_walks(rng, lens), interval=4.0, jitter=3.0, span=300.0)
    t = ts.t.copy()
    t[ts.offsets[6] + 3] = t[ts.offsets[6] + 2]  # a repeated timestamp
    t[ts.offsets[7]:ts.offsets[8]] += 1e5         # no common time with anything
    from trajsimimeasures_amd.data import TrajSet
    ts = TrajSet(ts.xy, ts.offsets, t)
    n = ts.n
    pa, pb = np.meshgrid(np.arange(n), np.arange(n))
    pa, pb = pa.ravel().astype(np.int32), pb.ravel().astype(np.int32)
    d = engine.DeviceTrajSet(ts)
    for M, R in ((10, 64), (3, 1), (100, 64), (5000, 1)):
        got = engine.pairs("sar", d, pa, pb, eps=0.01, time_step=5, target_length=M, replicas=R)
        got = got.cpu().numpy()
        _check(ts, pa, pb, got, 0.01, 5, M, R)
    # an identical pair with grid-aligned times: every sample matches
    xy = np.random.default_rng(66).random((200, 2))
    tg = TrajSet(xy, np.array([0, 200]), 10.0 * np.arange(200))
    dg = engine.DeviceTrajSet(tg)
    same = engine.pairs("sar", dg, [0], [0], eps=1e-9, time_step=10, target_length=500, replicas=64).cpu().numpy()
    assert same[0] == 64 * 200 * 10
    binned = engine.pairs("sar", dg, [0], [0], eps=1e-9, time_step=10, target_length=7, replicas=64).cpu().numpy()
    assert binned[0] == (200 + 7 - 1) * 10


def test_sar_grid_triu_dropin_and_bad_index(gpu):
    from trajsimimeasures_amd import core, engine
    from trajsimimeasures_amd.core import pad_lists_to_array
    ts, _, _ = _st_set(80, 5, 90, seed=67, span=300.0)
    d = engine.DeviceTrajSet(ts)
    n = ts.n
    kw = dict(eps=0.02, time_step=10, target_length=10, replicas=64)
    g = engine.grid("sar", d, 0, 25, 25, n - 25, **kw).cpu().numpy()
    qa, qb = np.repeat(np.arange(25), n - 25), np.tile(np.arange(25, n), 25)
    assert np.array_equal(g.ravel(), O.pairs_sar(ts.xy, ts.t, ts.offsets, qa, qb, 0.02, 10, 10, 64))
    tri = engine.triu("sar", d, 40, **kw).cpu().numpy()
    ia, ib = np.triu_indices(40, 1)
    assert np.array_equal(tri, O.pairs_sar(ts.xy, ts.t, ts.offsets, ia, ib, 0.02, 10, 10, 64))
    # padded [B, L, 3] drop-in, reference signature
    # sar(t1, len1, t2, len2, distance_eps, time_eps, target_length, threads_per_traj=64)
    trajs = [np.concatenate([ts.traj(i), ts.t[ts.offsets[i]:ts.offsets[i + 1], None]], 1) for i in range(n)]
    t1, t2 = pad_lists_to_array(trajs[:40]), pad_lists_to_array(trajs[40:80])
    l1, l2 = [len(t) for t in trajs[:40]], [len(t) for t in trajs[40:80]]
    for R in (64, 1):
        got = core.sar(t1, l1, t2, l2, 0.02, 10, 10, threads_per_traj=R)
        ref = O.pairs_sar(ts.xy, ts.t, ts.offsets, np.arange(40), np.arange(40, 80), 0.02, 10, 10, R)
        assert np.array_equal(got, ref)
    # an index outside the set gives NaN and leaves the others alone
    pa = np.array([0, n, 3, -1], np.int32)
    pb = np.array([1, 2, n + 5, 4], np.int32)
    out = engine.pairs("sar", d, pa, pb, validate=False, **kw).cpu().numpy()
    assert np.isnan(out[[1, 2, 3]]).all()
    assert out[0] == O.pairs_sar(ts.xy, ts.t, ts.offsets, [0], [1], 0.02, 10, 10, 64)[0]
