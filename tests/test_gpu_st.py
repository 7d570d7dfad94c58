"""Parity of the spatio-temporal measures STEDR and STLCSS (core/stedr.py,
core/stlcss.py) on the DP kernel against the FP64 oracle: match counts
bit-exact (hence the normalised values), for both eps conventions, edge
lengths, column strips, explicit pairs, grid and triangle addressing, and the
padded drop-in functions."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import uniform_pairs

pytestmark = pytest.mark.gpu


def _st_set(n, lo, hi, seed):
    from trajsimimeasures_amd.data import with_timestamps
    ts, pa, pb = uniform_pairs(n, lo=lo, hi=hi, seed=seed)
    return with_timestamps(np.random.default_rng(seed + 1), ts), pa, pb


def _check(measure, ts, pa, pb, got, cnt, eps, delta, eps_mode="sqrt"):
    ref, rcnt = O.pairs_st(measure, ts.xy, ts.t, ts.offsets, pa, pb, eps, delta, eps_mode)
    bad = np.nonzero(cnt != rcnt)[0]
    assert bad.size == 0, f"{measure}: {bad.size} count mismatches, first {bad[:5]}: {cnt[bad[:5]]} vs {rcnt[bad[:5]]}"
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("measure", ["stedr", "stlcss"])
@pytest.mark.parametrize("eps_mode", ["sqrt", "squared"])
def test_st_random_pairs(gpu, measure, eps_mode):
    from trajsimimeasures_amd import engine
    print(f"[case] st random pairs measure={measure} mode={eps_mode}", flush=True)
    ts, pa, pb = _st_set(2000, 5, 150, seed=41)
    d = engine.DeviceTrajSet(ts)
    # the pairs share the same start-time window, so time matches are common
    for eps, delta in ((0.0025, 60.0), (0.01, 30.0), (0.05, 600.0), (0.01, 0.0), (0.01, np.inf)):
        got, cnt = engine.pairs(measure, d, pa, pb, eps=eps, delta=delta, eps_mode=eps_mode, return_counts=True)
        _check(measure, ts, pa, pb, got.cpu().numpy(), cnt.cpu().numpy(), eps, delta, eps_mode)


def test_st_infinite_delta_is_edr_lcss(gpu):
    from trajsimimeasures_amd import engine
    ts, pa, pb = _st_set(1000, 5, 120, seed=43)
    d = engine.DeviceTrajSet(ts)
    for st, base in (("stedr", "edr"), ("stlcss", "lcss")):
        a = engine.pairs(st, d, pa, pb, eps=0.01, delta=np.inf).cpu().numpy()
        b = engine.pairs(base, d, pa, pb, eps=0.01).cpu().numpy()
        assert np.array_equal(a, b)


@pytest.mark.parametrize("measure", ["stedr", "stlcss"])
def test_st_edge_lengths_and_strips(gpu, measure):
    """Lengths 1 and 2, and pairs wider than 8 lanes x 16 columns (strips)."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import lengths_uniform, random_walks, with_timestamps
    print(f"[case] st edges measure={measure}", flush=True)
    rng = np.random.default_rng(45)
    lens = np.concatenate([[1, 2, 1, 2, 300, 1100, 129, 128, 17], lengths_uniform(rng, 40, 1, 400)])
    ts = with_timestamps(rng, random_walks(rng, lens), interval=2.0, jitter=1.0, span=20.0)
    n = ts.n
    pa, pb = np.meshgrid(np.arange(n), np.arange(n))
    pa, pb = pa.ravel().astype(np.int32), pb.ravel().astype(np.int32)
    d = engine.DeviceTrajSet(ts)
    got, cnt = engine.pairs(measure, d, pa, pb, eps=0.02, delta=5.0, return_counts=True)
    _check(measure, ts, pa, pb, got.cpu().numpy(), cnt.cpu().numpy(), 0.02, 5.0)


def test_st_grid_triu_and_dropin(gpu):
    from trajsimimeasures_amd import core, engine
    from trajsimimeasures_amd.core import pad_lists_to_array
    ts, _, _ = _st_set(60, 5, 80, seed=47)
    d = engine.DeviceTrajSet(ts)
    n = ts.n
    g = engine.grid("stlcss", d, 0, 20, 20, n - 20, eps=0.01, delta=45.0).cpu().numpy()
    qa, qb = np.repeat(np.arange(20), n - 20), np.tile(np.arange(20, n), 20)
    ref, _ = O.pairs_st("stlcss", ts.xy, ts.t, ts.offsets, qa, qb, 0.01, 45.0)
    assert np.array_equal(g.ravel(), ref)
    tri = engine.triu("stedr", d, 30, eps=0.01, delta=45.0).cpu().numpy()
    ia, ib = np.triu_indices(30, 1)
    ref, _ = O.pairs_st("stedr", ts.xy, ts.t, ts.offsets, ia, ib, 0.01, 45.0)
    assert np.array_equal(tri, ref)
    # padded [B, L, 3] drop-in, reference signature stedr(t1, len1, t2, len2, eps, delta)
    trajs = [np.concatenate([ts.traj(i), ts.t[ts.offsets[i]:ts.offsets[i + 1], None]], 1) for i in range(n)]
    t1 = pad_lists_to_array(trajs[:30])
    t2 = pad_lists_to_array(trajs[30:60])
    l1 = [len(t) for t in trajs[:30]]
    l2 = [len(t) for t in trajs[30:60]]
    for m, f in (("stedr", core.stedr), ("stlcss", core.stlcss)):
        got = f(t1, l1, t2, l2, 0.01, 45.0)
        ref, _ = O.pairs_st(m, ts.xy, ts.t, ts.offsets, np.arange(30), np.arange(30, 60), 0.01, 45.0)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("eps", [0.0025, 0.02, 0.2])
def test_cdds_matches_oracle(gpu, eps):
    """CDDS (core/cdds.py) on the GPU adds the same FP64 terms in the same
    order as the oracle: equal values; a pair of identical trajectories gives
    their duration."""
    from trajsimimeasures_amd import engine
    print(f"[case] cdds eps={eps}", flush=True)
    ts, pa, pb = _st_set(1500, 2, 150, seed=51)
    d = engine.DeviceTrajSet(ts)
    got = engine.pairs("cdds", d, pa, pb, eps=eps).cpu().numpy()
    ref = O.pairs_cdds(ts.xy, ts.t, ts.offsets, pa, pb, eps)
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    same = engine.pairs("cdds", d, pa[:50], pa[:50], eps=eps).cpu().numpy()
    dur = np.array([ts.t[ts.offsets[i + 1] - 1] - ts.t[ts.offsets[i]] for i in pa[:50]])
    assert np.allclose(same, dur, rtol=1e-12, atol=0)


def test_cdds_scrambled_and_repeated_timestamps(gpu):
    """Second trajectories with decreasing timestamps take the full walk,
    non-decreasing ones (repeated timestamps included) the time window; both
    equal the oracle."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import TrajSet
    print("[case] cdds scrambled / repeated timestamps", flush=True)
    ts, pa, pb = _st_set(1200, 2, 120, seed=55)
    rng = np.random.default_rng(56)
    t = ts.t.copy()
    for i in rng.choice(ts.n, ts.n // 3, replace=False):
        o, n = ts.offsets[i], ts.offsets[i + 1] - ts.offsets[i]
        a, b = o + rng.integers(0, n, 2)
        if rng.random() < 0.5:
            t[a], t[b] = t[b], t[a]       # decreasing somewhere
        elif n > 1:
            t[o + 1] = t[o]               # a zero-duration segment
    ts2 = TrajSet(ts.xy, ts.offsets, t)
    d = engine.DeviceTrajSet(ts2)
    for eps in (0.005, 0.05):
        got = engine.pairs("cdds", d, pa, pb, eps=eps).cpu().numpy()
        ref = O.pairs_cdds(ts2.xy, ts2.t, ts2.offsets, pa, pb, eps)
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


def test_cdds_dropin_and_long(gpu):
    from trajsimimeasures_amd import core
    from trajsimimeasures_amd.core import pad_lists_to_array
    from trajsimimeasures_amd.data import lengths_uniform, random_walks, with_timestamps
    rng = np.random.default_rng(53)
    lens = np.concatenate([[1, 2, 65, 64, 130, 700], lengths_uniform(rng, 34, 1, 90)])
    ts = with_timestamps(rng, random_walks(rng, lens), interval=5.0, jitter=4.0, span=60.0)
    trajs = [np.concatenate([ts.traj(i), ts.t[ts.offsets[i]:ts.offsets[i + 1], None]], 1) for i in range(ts.n)]
    t1, t2 = pad_lists_to_array(trajs[:20]), pad_lists_to_array(trajs[20:])
    l1, l2 = [len(t) for t in trajs[:20]], [len(t) for t in trajs[20:]]
    got = core.cdds(t1, l1, t2, l2, 0.02)
    ref = O.pairs_cdds(ts.xy, ts.t, ts.offsets, np.arange(20), np.arange(20, 40), 0.02)
    assert np.array_equal(got, ref)
