"""Direct DP launches (few pairs, one wavefront per pair, no planner) against
the planned launch and the FP64 oracle.

Bars (written in each assertion):
  - direct and planned launches of DTW, EDR, LCSS, Frechet, STEDR, STLCSS
    (FP32 and FP64 modes) are equal bit for bit: every cell follows the
    same recurrence in the same order whichever lanes hold it;
  - ERP's boundary sums are reduced over the segment's lanes, so the two
    launches may differ by an FP64 rounding of the sum: equal within 1e-6
    relative in FP32 mode, 1e-13 in FP64 mode;
  - against the oracle: FP32 max relative error <= 1e-5, counts bit-exact,
    FP64 mode within 1e-13.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import fixed_lengths, uniform_pairs

pytestmark = pytest.mark.gpu

EPS = 0.01
DELTA = 60.0


def _set_direct(n):
    from trajsimimeasures_amd import _lib
    _lib.check(_lib.lib().tsm_tune_set_direct_pairs(n))


def _run(measure, d, pa, pb, **kw):
    from trajsimimeasures_amd import engine
    counting = measure in ("edr", "lcss", "stedr", "stlcss")
    out = engine.pairs(measure, d, pa, pb, return_counts=measure in ("edr", "lcss"), **kw)
    if isinstance(out, tuple):
        return out[0].cpu().numpy(), out[1].cpu().numpy()
    return out.cpu().numpy(), (None if not counting else out.cpu().numpy())


def _both(measure, d, pa, pb, **kw):
    """(direct, planned) results of one call."""
    _set_direct(-1)
    direct = _run(measure, d, pa, pb, **kw)
    _set_direct(0)
    try:
        planned = _run(measure, d, pa, pb, **kw)
    finally:
        _set_direct(-1)
    return direct, planned


def _lens_every_w(wm: int, seed: int):
    """Pairs whose shorter side falls in every 64-lane W bucket (4 .. wm
    columns per lane) at and around its boundaries, plus short and unequal
    pairs."""
    rng = np.random.default_rng(seed)
    cols = [1, 2, 3, 63, 64, 65, 200, 255, 256, 257]
    for w in range(8, wm + 1, 4):
        cols += [64 * w - 1, 64 * w]
    cols += list(rng.integers(1, 64 * wm + 1, 24))
    la = [int(c) + int(rng.integers(0, 40)) for c in cols]
    lb = list(cols)
    swap = rng.random(len(la)) < 0.5  # the longer side on either trajectory
    la2 = [b if s else a for a, b, s in zip(la, lb, swap)]
    lb2 = [a if s else b for a, b, s in zip(la, lb, swap)]
    return la2, lb2


@pytest.mark.parametrize("measure", ["dtw", "edr", "lcss", "erp", "frechet"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_direct_equals_planned_every_w(gpu, measure, precision):
    """Sets up to 40 points longer than a direct launch admits (64 x 24): the
    default call is planned as well (its wide pairs on 64-lane segments,
    column strips beyond 1,536 columns), equal to the forced planned launch
    and to the oracle; test_direct_launch_every_w covers direct launches."""
    la, lb = _lens_every_w(24, seed=7)
    print(f"[case] direct vs planned measure={measure} precision={precision} pairs={len(la)}", flush=True)
    ts, pa, pb = fixed_lengths(la, lb, seed=8)
    from trajsimimeasures_amd import engine
    d = engine.DeviceTrajSet(ts)
    assert ts.lengths.max() <= 64 * 24 + 40
    g = ts.xy.mean(0)
    kw = dict(eps=EPS, gpoint=g, precision=precision)
    (dv, dc), (pv, pc) = _both(measure, d, pa, pb, **kw)
    if measure == "erp":
        tol = 1e-6 if precision == "fp32" else 1e-13
        assert np.max(np.abs(dv - pv) / np.maximum(np.abs(pv), 1e-300)) <= tol
    else:
        assert np.array_equal(dv, pv), (measure, np.nonzero(dv != pv)[0][:5])
    if dc is not None:
        assert np.array_equal(dc, pc)
    ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
    if measure in ("edr", "lcss"):
        assert np.array_equal(dc, rcnt) and np.array_equal(dv, ref)
    else:
        rtol = 1e-5 if precision == "fp32" else 1e-13
        rel = np.abs(dv - ref) / np.maximum(np.abs(ref), 1e-300)
        assert rel.max() <= rtol, (measure, precision, float(rel.max()))


@pytest.mark.parametrize("measure", ["stedr", "stlcss"])
def test_direct_equals_planned_st(gpu, measure):
    """STEDR/STLCSS (16 FP64 columns per lane): 64-lane W of 4 .. 16."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import with_timestamps
    la, lb = _lens_every_w(16, seed=9)
    print(f"[case] direct vs planned measure={measure} pairs={len(la)}", flush=True)
    ts, pa, pb = fixed_lengths(la, lb, seed=10)
    ts = with_timestamps(np.random.default_rng(11), ts)
    d = engine.DeviceTrajSet(ts)
    (dv, _), (pv, _) = _both(measure, d, pa, pb, eps=EPS, delta=DELTA)
    assert np.array_equal(dv, pv)
    ref, _ = O.pairs_st(measure, ts.xy, ts.t, ts.offsets, pa, pb, EPS, DELTA)
    assert np.array_equal(dv, ref)


@pytest.mark.parametrize("npairs", [1, 64, 512, 2048, 6144])
def test_direct_batch_sizes(gpu, npairs):
    """Batches up to the default direct limit (6,144 pairs on MI355X), DTW,
    EDR, ERP and Frechet against the oracle."""
    from trajsimimeasures_amd import engine
    print(f"[case] direct batch npairs={npairs}", flush=True)
    ts, pa, pb = uniform_pairs(npairs, seed=40 + npairs)
    d = engine.DeviceTrajSet(ts)
    g = ts.xy.mean(0)
    for m in ("dtw", "edr", "erp", "frechet"):
        got, cnt = _run(m, d, pa, pb, eps=EPS, gpoint=g)
        ref, rcnt = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
        if m == "edr":
            assert np.array_equal(cnt, rcnt)
        else:
            assert np.max(np.abs(got - ref) / ref) <= 1e-5, m


def test_direct_bad_indices(gpu):
    """Pair indices outside the set: NaN and count -1 in a direct launch,
    as in a planned one; the other pairs are computed."""
    from trajsimimeasures_amd import engine
    ts, pa, pb = uniform_pairs(40, seed=12)
    d = engine.DeviceTrajSet(ts)
    pa = pa.copy()
    pb = pb.copy()
    pa[3] = -1
    pb[7] = ts.n
    pa[11] = ts.n + 5
    bad = np.zeros(pa.size, bool)
    bad[[3, 7, 11]] = True
    for direct in (-1, 0):
        _set_direct(direct)
        try:
            v, c = _run("edr", d, pa, pb, eps=EPS, validate=False)
        finally:
            _set_direct(-1)
        assert np.all(np.isnan(v[bad])) and np.all(c[bad] == -1), direct
        ref, rcnt = O.pairs("edr", ts.xy, ts.offsets, pa[~bad], pb[~bad], eps=EPS)
        assert np.array_equal(c[~bad], rcnt), direct


def test_direct_long_set_takes_planner(gpu):
    """A set longer than 64 x 24 points needs column strips, which only the
    planned launch has: a small call on it is planned, and correct."""
    from trajsimimeasures_amd import engine
    ts, pa, pb = fixed_lengths([1700, 30, 200], [1600, 40, 1650], seed=13)
    d = engine.DeviceTrajSet(ts)
    got, _ = _run("dtw", d, pa, pb)
    ref, _ = O.pairs("dtw", ts.xy, ts.offsets, pa, pb)
    assert np.max(np.abs(got - ref) / ref) <= 1e-5


def test_direct_graph_replay(gpu):
    """A captured direct launch (512 pairs) replays bit-identical to eager."""
    import torch

    from trajsimimeasures_amd import engine
    ts, pa, pb = uniform_pairs(512, seed=14)
    d = engine.DeviceTrajSet(ts)
    pa_dev = torch.as_tensor(pa, dtype=torch.int32, device="cuda")
    pb_dev = torch.as_tensor(pb, dtype=torch.int32, device="cuda")
    eager = _run("dtw", d, pa, pb)[0]
    g = engine.PairsGraph("dtw", d, pa_dev, pb_dev)
    for rep in range(3):
        g.out.fill_(-1.0)
        g.stream.synchronize()
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        assert np.array_equal(g.out.cpu().numpy(), eager), rep
        x = torch.randn(1 << 18, device="cuda").sum()
        torch.cuda.synchronize()
        del x
    g.close()


@pytest.mark.parametrize("measure", ["dtw", "edr", "erp"])
def test_direct_grid_and_triangle(gpu, measure):
    """Grid (7 x 40) and triangle (n = 50) addressing decoded in a direct
    launch: equal to the planned launch and to the oracle."""
    from trajsimimeasures_amd import engine
    from tests.workloads import lognormal_set
    ts = lognormal_set(90, seed=15, lo=5, hi=300)
    d = engine.DeviceTrajSet(ts)
    g = ts.xy.mean(0)
    counted = measure == "edr"
    outs = {}
    for direct in (-1, 0):
        _set_direct(direct)
        try:
            gr = engine.grid(measure, d, 0, 7, 7, 40, eps=EPS, gpoint=g, return_counts=counted)
            tr = engine.triu(measure, d, 50, eps=EPS, gpoint=g, return_counts=counted)
        finally:
            _set_direct(-1)
        if counted:
            gr, tr = gr[1], tr[1]
        outs[direct] = (gr.cpu().numpy().reshape(-1), tr.cpu().numpy().reshape(-1))
    for a, b in zip(outs[-1], outs[0]):
        if measure == "erp":
            assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) <= 1e-6
        else:
            assert np.array_equal(a, b)
    qa = np.repeat(np.arange(7), 40)
    qb = np.tile(np.arange(7, 47), 7)
    ia, ib = np.triu_indices(50, 1)
    for (pa, pb), got in (((qa, qb), outs[-1][0]), ((ia, ib), outs[-1][1])):
        ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
        if counted:
            assert np.array_equal(got, rcnt)
        else:
            assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5


@pytest.mark.parametrize("seed", range(8))
def test_random_launch_shapes(gpu, seed):
    """Random batch sizes (1 .. 12,000 pairs: direct and planned launches)
    and lengths (lognormal, median 80, up to 1,700 points: 64-lane buckets,
    column strips) for a random DP measure, against the oracle."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import random_walks
    rng = np.random.default_rng(100 + seed)
    npairs = int(np.exp(rng.uniform(0, np.log(12000))))
    measure = ["dtw", "edr", "lcss", "erp", "frechet"][seed % 5]
    lens = np.clip(np.round(np.exp(rng.normal(np.log(80), 1.0, 2 * npairs))), 1, 1700).astype(np.int64)
    ts = random_walks(rng, lens)
    pa = np.arange(0, 2 * npairs, 2)
    pb = pa + 1
    print(f"[case] random shapes seed={seed} measure={measure} pairs={npairs} max_len={lens.max()}", flush=True)
    d = engine.DeviceTrajSet(ts)
    g = ts.xy.mean(0)
    got, cnt = _run(measure, d, pa, pb, eps=EPS, gpoint=g)
    ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
    if measure in ("edr", "lcss"):
        assert np.array_equal(cnt, rcnt)
    else:
        assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5


@pytest.mark.parametrize("measure", ["dtw", "edr", "lcss", "erp", "frechet", "stedr", "stlcss"])
def test_long_row_buckets(gpu, measure):
    """Planned launches of up to 2^20 pairs put pairs with more than 384 rows
    on 16 lanes (up to 128 columns) or a whole wavefront (up to 384): pairs of
    385 .. 1,200 rows with columns at every boundary of that rule, against
    the oracle and, bit for bit (ERP within its boundary-sum rounding),
    against the same call with a 16-lane hint, which keeps the column-count
    buckets."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import with_timestamps
    rng = np.random.default_rng(31)
    cols = [1, 2, 15, 16, 17, 31, 100, 127, 128, 129, 200, 255, 256, 257, 300, 383, 384, 385, 500]
    rows = [385, 386, 400, 700, 1200]
    la = [int(r) for r in rows for _ in cols] + [384, 384, 60]
    lb = [int(c) for _ in rows for c in cols] + [30, 300, 1000]
    swap = rng.random(len(la)) < 0.5  # the longer side on either trajectory
    la, lb = [b if s else a for a, b, s in zip(la, lb, swap)], [a if s else b for a, b, s in zip(la, lb, swap)]
    print(f"[case] long-row buckets measure={measure} pairs={len(la)}", flush=True)
    ts, pa, pb = fixed_lengths(la, lb, seed=32)
    st = measure in ("stedr", "stlcss")
    if st:
        ts = with_timestamps(np.random.default_rng(33), ts)
    d = engine.DeviceTrajSet(ts)
    g = ts.xy.mean(0)
    kw = dict(eps=EPS, delta=DELTA) if st else dict(eps=EPS, gpoint=g)
    _set_direct(0)
    try:
        got, cnt = _run(measure, d, pa, pb, **kw)
        hint, hcnt = _run(measure, d, pa, pb, lanes=16, **kw)
    finally:
        _set_direct(-1)
    if measure == "erp":
        assert np.max(np.abs(got - hint) / np.maximum(np.abs(hint), 1e-300)) <= 1e-6
    else:
        assert np.array_equal(got, hint), (measure, np.nonzero(got != hint)[0][:5])
    if st:
        ref, _ = O.pairs_st(measure, ts.xy, ts.t, ts.offsets, pa, pb, EPS, DELTA)
        assert np.array_equal(got, ref)
    else:
        ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
        if measure in ("edr", "lcss"):
            assert np.array_equal(cnt, rcnt)
        else:
            assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5


def _lens_direct(wm: int, seed: int):
    """Like _lens_every_w, with every length within a direct launch's limit
    (64 wm points), so the call is a direct launch: every 64-lane W, rows up
    to the limit (the row trajectory staged in LDS)."""
    la, lb = _lens_every_w(wm, seed)
    cap = 64 * wm
    return [min(x, cap) for x in la], [min(x, cap) for x in lb]


@pytest.mark.parametrize("measure", ["dtw", "edr", "lcss", "erp", "frechet"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_direct_launch_every_w(gpu, measure, precision):
    """Direct launches (lengths up to 64 x 24) in every 64-lane W, rows up to
    1,536 points read from LDS: equal to the planned launch (ERP within its
    boundary-sum rounding) and to the oracle."""
    from trajsimimeasures_amd import engine
    la, lb = _lens_direct(24, seed=17)
    print(f"[case] direct every W measure={measure} precision={precision} pairs={len(la)}", flush=True)
    ts, pa, pb = fixed_lengths(la, lb, seed=18)
    assert ts.lengths.max() <= 64 * 24
    d = engine.DeviceTrajSet(ts)
    g = ts.xy.mean(0)
    kw = dict(eps=EPS, gpoint=g, precision=precision)
    (dv, dc), (pv, pc) = _both(measure, d, pa, pb, **kw)
    if measure == "erp":
        tol = 1e-6 if precision == "fp32" else 1e-13
        assert np.max(np.abs(dv - pv) / np.maximum(np.abs(pv), 1e-300)) <= tol
    else:
        assert np.array_equal(dv, pv), (measure, np.nonzero(dv != pv)[0][:5])
    if dc is not None:
        assert np.array_equal(dc, pc)
    ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
    if measure in ("edr", "lcss"):
        assert np.array_equal(dc, rcnt) and np.array_equal(dv, ref)
    else:
        rtol = 1e-5 if precision == "fp32" else 1e-13
        rel = np.abs(dv - ref) / np.maximum(np.abs(ref), 1e-300)
        assert rel.max() <= rtol, (measure, precision, float(rel.max()))


@pytest.mark.parametrize("measure", ["stedr", "stlcss"])
def test_direct_launch_st(gpu, measure):
    """STEDR/STLCSS direct launches (lengths up to 64 x 16, points and
    timestamps staged in LDS) in every 64-lane W: equal to the planned
    launch and to the oracle."""
    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import with_timestamps
    la, lb = _lens_direct(16, seed=19)
    print(f"[case] direct ST measure={measure} pairs={len(la)}", flush=True)
    ts, pa, pb = fixed_lengths(la, lb, seed=20)
    ts = with_timestamps(np.random.default_rng(21), ts)
    assert ts.lengths.max() <= 64 * 16
    d = engine.DeviceTrajSet(ts)
    (dv, _), (pv, _) = _both(measure, d, pa, pb, eps=EPS, delta=DELTA)
    assert np.array_equal(dv, pv)
    ref, _ = O.pairs_st(measure, ts.xy, ts.t, ts.offsets, pa, pb, EPS, DELTA)
    assert np.array_equal(dv, ref)
