"""Parity of the HIP kernels (through libtsm.so's C ABI) against the FP64
oracle.

Bars (written in each assertion):
  - DTW, ERP, discrete Frechet, Hausdorff in FP32: max relative error
    <= 1e-5 against the FP64 oracle on FP64 inputs;
  - EDR, LCSS: integer counts bit-exact (and hence the normalised values);
  - FP64 mode (DTW, ERP, Frechet): equal to the oracle within 1e-13 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import EPS, fixed_lengths, lognormal_set, uniform_pairs

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-5
FP64_RTOL = 1e-13


def _engine():
    from trajsimimeasures_amd import engine
    return engine


def _dev(ts):
    return _engine().DeviceTrajSet(ts)


def _gpu(measure, dset, pa, pb, **kw):
    out = _engine().pairs(measure, dset, pa, pb, return_counts=measure in ("edr", "lcss"), **kw)
    if isinstance(out, tuple):
        v, c = out
        return v.cpu().numpy(), c.cpu().numpy()
    return out.cpu().numpy(), None


def _check(measure, ts, pa, pb, got, cnt, *, eps=EPS, eps_mode="sqrt", gpoint=(0.0, 0.0), rtol=FP32_RTOL):
    ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=eps, eps_mode=eps_mode, gpoint=gpoint)
    if measure in ("edr", "lcss"):
        bad = np.nonzero(cnt != rcnt)[0]
        assert bad.size == 0, f"{measure}: {bad.size} count mismatches, first pair {bad[:5]}: {cnt[bad[:5]]} vs {rcnt[bad[:5]]}"
        assert np.array_equal(got, ref)
    else:
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        rel[(got == ref)] = 0.0
        k = int(np.argmax(rel))
        assert rel[k] <= rtol, f"{measure}: max rel err {rel[k]:.3e} at pair {k}: {got[k]!r} vs {ref[k]!r}"
    return ref


ALL = ["dtw", "edr", "lcss", "erp", "frechet", "hausdorff"]


@pytest.mark.parametrize("measure", ALL)
@pytest.mark.parametrize("lanes", [0, 4, 8])
def test_random_walk_pairs(gpu, measure, lanes):
    print(f"[case] random walk pairs measure={measure} lanes={lanes}", flush=True)
    ts, pa, pb = uniform_pairs(3000, seed=11)
    d = _dev(ts)
    g = ts.xy.mean(0)
    got, cnt = _gpu(measure, d, pa, pb, eps=EPS, gpoint=g, lanes=lanes)
    _check(measure, ts, pa, pb, got, cnt, gpoint=g)


@pytest.mark.parametrize("measure", ALL)
def test_edge_lengths(gpu, measure):
    """Length 1 on either side, equal lengths, column counts around the
    lane/register boundaries (L*W), orientation swaps."""
    print(f"[case] edge lengths measure={measure}", flush=True)
    la = [1, 1, 2, 1, 5, 16, 17, 15, 8, 9, 33, 64, 65, 127, 128, 129, 256, 257, 3, 200, 1, 40]
    lb = [1, 7, 1, 300, 5, 16, 16, 17, 9, 8, 40, 64, 66, 128, 128, 130, 256, 260, 190, 3, 2, 40]
    ts, pa, pb = fixed_lengths(la, lb, seed=3)
    d = _dev(ts)
    for lanes in (0, 4, 8):
        got, cnt = _gpu(measure, d, pa, pb, eps=0.02, lanes=lanes)
        _check(measure, ts, pa, pb, got, cnt, eps=0.02)


@pytest.mark.parametrize("wide", [0, 16, 64])
@pytest.mark.parametrize("measure", ["dtw", "edr", "lcss", "erp", "frechet"])
def test_column_strips(gpu, measure, wide):
    """Pairs wider than one strip of L*W register columns (swept in strips
    with a boundary column in scratch): up to 2000 x 1999, with the pairs
    wider than 384 columns on 16-lane strips (wide = 16) or whole
    wavefronts (64; 0 picks 64 for a launch this small), across every
    column split of the 64-lane buckets (W = 12, 16, 20, 24) and past one
    1536-column strip."""
    from trajsimimeasures_amd import _lib
    print(f"[case] column strips measure={measure} wide={wide}", flush=True)
    la = [300, 420, 600, 1200, 257, 385, 1000, 193, 192, 400, 2000, 769, 768, 1537, 3100, 1300, 1025]
    lb = [290, 500, 610, 1100, 300, 384, 999, 195, 192, 768, 1999, 800, 770, 1536, 3090, 1290, 1030]
    ts, pa, pb = fixed_lengths(la, lb, seed=5)
    d = _dev(ts)
    g = (0.5, 0.5)
    _lib.check(_lib.lib().tsm_tune_set_wide_lanes(wide))
    try:
        for lanes in (0, 4, 8):
            got, cnt = _gpu(measure, d, pa, pb, eps=0.01, gpoint=g, lanes=lanes)
            _check(measure, ts, pa, pb, got, cnt, eps=0.01, gpoint=g)
    finally:
        _lib.check(_lib.lib().tsm_tune_set_wide_lanes(0))


@pytest.mark.parametrize("measure", ["edr", "lcss"])
@pytest.mark.parametrize("eps_mode", ["sqrt", "squared"])
def test_eps_conventions(gpu, measure, eps_mode):
    """Both match conventions, bit-exact; includes points placed exactly at
    distance eps so the two conventions can disagree."""
    print(f"[case] eps conventions measure={measure} mode={eps_mode}", flush=True)
    ts, pa, pb = uniform_pairs(500, lo=5, hi=60, seed=21)
    d = _dev(ts)
    for eps in (0.0, 0.001, 0.0025, 0.01, 0.1, np.inf):
        got, cnt = _gpu(measure, d, pa, pb, eps=eps, eps_mode=eps_mode)
        _check(measure, ts, pa, pb, got, cnt, eps=eps, eps_mode=eps_mode)


def test_eps_boundary_points(gpu):
    """Points at distance exactly eps apart along an axis: the sqrt and
    squared conventions and the oracle must agree cell for cell."""
    from trajsimimeasures_amd.data import pack_list
    eps = 0.0025
    a = np.stack([np.linspace(0.1, 0.3, 50), np.full(50, 0.5)], 1)
    b = a + np.array([0.0, eps])
    c = a + np.array([eps, 0.0])
    ts = pack_list([a, b, c, a * 1.0000001])
    pa = np.array([0, 0, 1, 3])
    pb = np.array([1, 2, 2, 0])
    dset = _dev(ts)
    for mode in ("sqrt", "squared"):
        for m in ("edr", "lcss"):
            got, cnt = _gpu(m, dset, pa, pb, eps=eps, eps_mode=mode)
            _check(m, ts, pa, pb, got, cnt, eps=eps, eps_mode=mode)


@pytest.mark.parametrize("shift", [0.0, 5.0, 1000.0])
def test_count_match_paths(gpu, shift):
    """EDR and LCSS test matches on the FP32 copy against a guard band when
    the set's coordinate bound allows it and in FP64 otherwise (coord_max =
    0).  The FP32 copy is centred, so shifting the set (5, 1000) leaves the
    band as narrow as at 0.  Every path is bit-exact against the oracle."""
    from trajsimimeasures_amd.data import TrajSet
    print(f"[case] count match paths shift={shift}", flush=True)
    ts, pa, pb = uniform_pairs(2000, lo=5, hi=120, seed=31)
    ts = TrajSet(ts.xy + shift, ts.offsets)
    d = _dev(ts)
    for m in ("edr", "lcss"):
        for eps in (0.0025, 0.01):
            got, cnt = _gpu(m, d, pa, pb, eps=eps)
            _check(m, ts, pa, pb, got, cnt, eps=eps)
            cm, d.coord_max = d.coord_max, 0.0
            got64, cnt64 = _gpu(m, d, pa, pb, eps=eps)
            d.coord_max = cm
            assert np.array_equal(cnt, cnt64) and np.array_equal(got, got64)


@pytest.mark.parametrize("measure", ["dtw", "erp", "frechet", "hausdorff"])
def test_fp64_mode(gpu, measure):
    print(f"[case] fp64 measure={measure}", flush=True)
    ts, pa, pb = fixed_lengths([20, 150, 300, 1100, 1, 33], [45, 150, 310, 1050, 9, 1], seed=8)
    d = _dev(ts)
    g = (0.3, 0.7)
    got, _ = _gpu(measure, d, pa, pb, gpoint=g, precision="fp64")
    _check(measure, ts, pa, pb, got, None, gpoint=g, rtol=FP64_RTOL)


def test_identity_exact_zero(gpu):
    ts, pa, _ = uniform_pairs(300, seed=4)
    d = _dev(ts)
    for m in ALL:
        got, cnt = _gpu(m, d, pa, pa, eps=EPS)
        # Hausdorff too: a trajectory's last point projects onto its own last
        # segment at u >= 1 exactly (1/|s|^2 is rounded up), residual 0
        assert np.all(got == 0.0), m


def test_swap_symmetry(gpu):
    """Swapping t1/t2 of pairs with unequal lengths runs the same program on
    the same orientation: bit-identical results."""
    ts, pa, pb = uniform_pairs(1000, seed=6)
    keep = ts.lengths[pa] != ts.lengths[pb]
    pa, pb = pa[keep], pb[keep]
    d = _dev(ts)
    for m in ALL:
        x, _ = _gpu(m, d, pa, pb, eps=EPS)
        y, _ = _gpu(m, d, pb, pa, eps=EPS)
        assert np.array_equal(x, y), m


@pytest.mark.parametrize("measure", ["dtw", "edr", "hausdorff", "frechet"])
def test_grid_and_triu(gpu, measure):
    """Q x D grid and upper-triangle addressing agree with explicit pairs."""
    from trajsimimeasures_amd import engine
    ts = lognormal_set(90, seed=2, hi=400)
    d = _dev(ts)
    q0, nq, d0, nd = 0, 12, 12, 78
    gq = engine.grid(measure, d, q0, nq, d0, nd, eps=EPS).cpu().numpy()
    qa = np.repeat(np.arange(q0, q0 + nq), nd)
    qb = np.tile(np.arange(d0, d0 + nd), nq)
    ref, rc = O.pairs(measure, ts.xy, ts.offsets, qa, qb, eps=EPS)
    _check(measure, ts, qa, qb, gq.reshape(-1), rc if measure == "edr" else None, eps=EPS)
    n = 60
    tri = engine.triu(measure, d, n, eps=EPS).cpu().numpy()
    iu = np.triu_indices(n, 1)
    ref2, rc2 = O.pairs(measure, ts.xy, ts.offsets, iu[0], iu[1], eps=EPS)
    _check(measure, ts, iu[0], iu[1], tri, rc2 if measure == "edr" else None, eps=EPS)


def test_dropin_core_padded(gpu):
    """The reference call surface: padded batches in, float64 [B] out."""
    from trajsimimeasures_amd import core
    ts, pa, pb = uniform_pairs(256, seed=9)
    L = int(ts.lengths.max())
    B = pa.size
    t1 = np.zeros((B, L, 2))
    t2 = np.zeros((B, L, 2))
    len1 = ts.lengths[pa]
    len2 = ts.lengths[pb]
    for k in range(B):
        t1[k, :len1[k]] = ts.traj(pa[k])
        t2[k, :len2[k]] = ts.traj(pb[k])
    g = np.array([0.4, 0.6])
    calls = {
        "dtw": core.dtw(t1, len1, t2, len2),
        "edr": core.edr(t1, len1, t2, list(len2), EPS),
        "lcss": core.lcss(t1, list(len1), t2, len2, EPS, threads_per_traj=16),
        "erp": core.erp(t1, len1, t2, len2, gpoint=g),
        "frechet": core.dfrechet(t1, len1, t2, len2, threads_per_traj=8),
        "hausdorff": core.hausdorff(t1, len1, t2, len2),
    }
    for m, got in calls.items():
        assert got.dtype == np.float64 and got.shape == (B,)
        ref, _ = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=EPS, gpoint=g)
        if m in ("edr", "lcss"):
            assert np.array_equal(got, ref), m
        else:
            assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= FP32_RTOL, m


def test_full_size_six_measures(gpu):
    """The 100,000-pair workload (lengths uniform 20..200, eps 0.0025, ERP gap
    point = centroid) for all six measures against the oracle run on all host
    cores."""
    print("[case] full 100k pairs, six measures", flush=True)
    ts, pa, pb = uniform_pairs(100_000, seed=0)
    d = _dev(ts)
    g = ts.xy.mean(0)
    for m in ALL:
        got, cnt = _gpu(m, d, pa, pb, eps=EPS, gpoint=g)
        _check(m, ts, pa, pb, got, cnt, gpoint=g)
        print(f"  {m} ok", flush=True)


def test_out_of_range_pairs_give_nan(gpu):
    """A pair index outside [0, n) (caller error that skipped validation)
    yields NaN (count -1) for that pair only; the kernels check the indices
    before any access."""
    from trajsimimeasures_amd.data import with_timestamps
    ts, pa, pb = uniform_pairs(300, lo=5, hi=80, seed=61)
    ts = with_timestamps(np.random.default_rng(62), ts)
    n = ts.n
    pa = pa.astype(np.int32).copy()
    pb = pb.astype(np.int32).copy()
    bad = np.array([0, 7, 150, 299])
    pa[bad[:2]] = [n, -1]
    pb[bad[2:]] = [n + 1000, -5]
    good = np.setdiff1d(np.arange(pa.size), bad)
    d = _dev(ts)
    for m in ALL + ["stedr", "cdds"]:
        print(f"[case] out-of-range pairs measure={m}", flush=True)
        out = _engine().pairs(m, d, pa, pb, eps=EPS, delta=60.0, validate=False,
                              return_counts=m in ("edr", "lcss", "stedr"))
        got, cnt = (out[0].cpu().numpy(), out[1].cpu().numpy()) if isinstance(out, tuple) else (out.cpu().numpy(), None)
        assert np.all(np.isnan(got[bad])), m
        assert not np.any(np.isnan(got[good])), m
        if cnt is not None:
            assert np.all(cnt[bad] == -1), m
        if m in ALL:
            ref, _ = O.pairs(m, ts.xy, ts.offsets, pa[good], pb[good], eps=EPS)
            if m in ("edr", "lcss"):
                assert np.array_equal(got[good], ref)
            else:
                assert np.max(np.abs(got[good] - ref) / np.maximum(ref, 1e-300)) <= FP32_RTOL
