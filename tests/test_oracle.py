"""Pins the FP64 oracle (oracle/tsm_oracle.c) against what the mathematics
fixes: brute-force enumeration over explicit alignments, closed forms, hand
cases that separate the reference's boundary variants from the textbook
ones, invariants, and the EDR/LCSS threshold equivalence.  The reference
ships no golden vectors for this path (its "tests" are timing drivers), so
these checks are the oracle's anchor."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import bruteforce as BF


def rnd(rng, n, scale=1.0):
    return rng.random((n, 2)) * scale


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(1234)


# ---------------------------------------------------------------- brute force

@pytest.mark.parametrize("n1,n2", [(1, 1), (1, 4), (3, 1), (2, 3), (4, 4), (5, 3)])
def test_dtw_frechet_bruteforce(rng, n1, n2):
    for _ in range(5):
        a, b = rnd(rng, n1), rnd(rng, n2)
        assert O.single("dtw", a, b) == pytest.approx(BF.dtw(a, b), rel=1e-12, abs=1e-15)
        assert O.single("frechet", a, b) == BF.frechet(a, b)  # a max/min of the same distances: exact


@pytest.mark.parametrize("n1,n2", [(1, 1), (1, 3), (3, 2), (3, 3), (4, 3)])
def test_edr_lcss_bruteforce(rng, n1, n2):
    for eps in (0.05, 0.3, 0.6):
        thr = O.eps_threshold(eps)
        match = lambda p, q: (p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 < thr
        for _ in range(4):
            a, b = rnd(rng, n1), rnd(rng, n2)
            assert O.count("edr", a, b, eps) == BF.edr_count(a, b, match)
            assert O.count("lcss", a, b, eps) == BF.lcss_count(a, b, match)


@pytest.mark.parametrize("n1,n2", [(1, 1), (2, 1), (1, 3), (3, 3), (4, 2)])
def test_erp_bruteforce(rng, n1, n2):
    for _ in range(4):
        a, b = rnd(rng, n1), rnd(rng, n2)
        g = rng.random(2)
        assert O.single("erp", a, b, gpoint=g) == pytest.approx(BF.erp(a, b, g), rel=1e-12)


def test_hausdorff_textbook(rng):
    # agrees with the textbook point-to-segment Hausdorff whenever no
    # projection parameter lands in (0, 1e-5), where the reference takes the
    # nearer endpoint instead
    for n1, n2 in [(2, 2), (3, 5), (6, 4), (1, 4), (4, 1), (1, 1)]:
        for _ in range(5):
            a, b = rnd(rng, n1), rnd(rng, n2)
            assert O.single("hausdorff", a, b) == pytest.approx(BF.hausdorff(a, b), rel=1e-12, abs=1e-15)


# ---------------------------------------------------------------- closed forms / hand cases

def test_dtw_one_point_closed_form(rng):
    p = rng.random((1, 2))
    b = rnd(rng, 7)
    s = 0.0
    for q in b:  # D[1][j] = d_j + D[1][j-1]: the in-order running sum, bit for bit
        s += math.sqrt((p[0, 0] - q[0]) ** 2 + (p[0, 1] - q[1]) ** 2)
    assert O.single("dtw", p, b) == s
    assert O.single("frechet", p, b) == max(math.sqrt((p[0, 0] - q[0]) ** 2 + (p[0, 1] - q[1]) ** 2) for q in b)


def test_dtw_hand_case():
    a = [(0, 0), (1, 0), (2, 0)]
    b = [(0, 0), (2, 0)]
    assert O.single("dtw", a, b) == 1.0
    assert O.single("frechet", [(0, 0), (1, 0)], [(0, 1), (1, 1)]) == 1.0


def test_identity_and_symmetry(rng):
    a, b = rnd(rng, 9), rnd(rng, 6)
    for m in ("dtw", "frechet", "erp", "hausdorff"):
        assert O.single(m, a, a) == 0.0
        assert O.single(m, a, b) == O.single(m, b, a)
    for m in ("edr", "lcss"):
        assert O.single(m, a, a, eps=1e-9) == 0.0
        assert O.count(m, a, b, 0.2) == O.count(m, b, a, 0.2)


def test_edr_lcss_extreme_eps(rng):
    a, b = rnd(rng, 5), rnd(rng, 8)
    # eps <= 0 never matches: zero-boundary EDR count = min(n1, n2) (the
    # textbook boundary would give max); LCSS count = 0
    assert O.count("edr", a, b, 0.0) == 5
    assert O.single("edr", a, b, eps=0.0) == 5 / 8
    assert O.single("lcss", a, b, eps=-1.0) == 1.0
    # eps = inf always matches: EDR 0 (textbook would give |n1 - n2| / max), LCSS min(n1, n2)
    assert O.single("edr", a, b, eps=math.inf) == 0.0
    assert O.count("lcss", a, b, math.inf) == 5
    assert O.single("lcss", a, b, eps=math.inf) == 0.0


def test_edr_zero_boundary_hand_case():
    # the second point of b matches a's only point after skipping b[0] for free
    assert O.single("edr", [(5, 5)], [(0, 0), (5, 5)], eps=1.0) == 0.0


def test_erp_closed_forms():
    # one point each: min(ga + gb, d) = d by the triangle inequality
    assert O.single("erp", [(0.3, 0.4)], [(0.6, 0.8)], gpoint=(0, 0)) == pytest.approx(0.5, rel=1e-15)
    # constant-sum boundary D[i][0] = Sa: 3.0 (the cumulative boundary gives 0.0)
    assert O.single("erp", [(0, 0), (3, 0)], [(3, 0)], gpoint=(0, 0)) == 3.0


def test_erp_translation_invariance(rng):
    a, b, g = rnd(rng, 6), rnd(rng, 4), rng.random(2)
    t = np.array([0.25, -0.5])
    assert O.single("erp", a + t, b + t, gpoint=g + t) == pytest.approx(O.single("erp", a, b, gpoint=g), rel=1e-12)


def test_hausdorff_hand_cases():
    # point-to-segment: 2.0 (point-to-point would give sqrt(5))
    assert O.single("hausdorff", [(0, 0), (2, 0)], [(1, 1), (1, 2)]) == 2.0
    # a projection parameter in (0, 1e-5) takes the nearer endpoint
    p, s = [(5e-6, 1.0)], [(0.0, 0.0), (1.0, 0.0)]
    assert O.single("hausdorff", p, s) == math.sqrt(5e-6 ** 2 + 1.0)
    # a degenerate segment uses the distance to its first endpoint
    assert O.single("hausdorff", [(0.0, 0.0)], [(3.0, 4.0), (3.0, 4.0)]) == 5.0
    # one-point polylines contribute 0 (core/hausdorff.py:50-51)
    assert O.single("hausdorff", [(0.0, 0.0)], [(3.0, 4.0)]) == 0.0


# ---------------------------------------------------------------- epsilon threshold

def test_eps_threshold_equivalence(rng):
    for eps in np.concatenate([rng.random(200) * 0.01, [0.0025, 1e-300, 1.0, 3.0]]):
        thr = O.eps_threshold(eps, "sqrt")
        cands = [thr, np.nextafter(thr, 0), np.nextafter(thr, np.inf), eps * eps,
                 np.nextafter(eps * eps, 0), np.nextafter(eps * eps, np.inf)]
        for d2 in cands:
            assert (d2 < thr) == (math.sqrt(d2) < eps), (eps, d2)
        assert O.eps_threshold(eps, "squared") == eps * eps


def test_eps_threshold_differs_from_square(rng):
    # the sqrt convention is not the same test as d2 < eps*eps for every eps
    eps = rng.random(2000) * 0.01
    assert any(O.eps_threshold(e, "sqrt") != e * e for e in eps)


# ---------------------------------------------------------------- batched entry

def test_pairs_matches_single(rng):
    from trajsimimeasures_amd.data import lengths_uniform, random_walks
    ts = random_walks(rng, lengths_uniform(rng, 40, 1, 30))
    pa = np.arange(0, 40, 2)
    pb = pa + 1
    g = ts.xy.mean(0)
    for m in ("dtw", "edr", "lcss", "erp", "frechet", "hausdorff"):
        vals, cnt = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=0.01, gpoint=g, threads=3)
        vals1, _ = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=0.01, gpoint=g, threads=1)
        assert np.array_equal(vals, vals1)
        for k in range(pa.size):
            ref = O.single(m, ts.traj(pa[k]), ts.traj(pb[k]), eps=0.01, gpoint=g)
            assert vals[k] == ref
            if m in ("edr", "lcss"):
                assert cnt[k] == O.count(m, ts.traj(pa[k]), ts.traj(pb[k]), 0.01)


def _edr_short_rule(a, b, thr):
    """EDR count with the update D = diag if match else min3(diag, up, left) + 1,
    which the engine uses; equal to the reference recurrence because unit-cost
    edit distances satisfy diag <= min(up, left) + 1."""
    n1, n2 = len(a), len(b)
    D = np.zeros((n1 + 1, n2 + 1), dtype=np.int64)
    for i in range(1, n1 + 1):
        for j in range(1, n2 + 1):
            d2 = (a[i - 1, 0] - b[j - 1, 0]) ** 2 + (a[i - 1, 1] - b[j - 1, 1]) ** 2
            D[i, j] = D[i - 1, j - 1] if d2 < thr else min(D[i - 1, j - 1], D[i - 1, j], D[i, j - 1]) + 1
    return int(D[n1, n2])


def test_edr_short_update_rule(rng):
    for _ in range(300):
        n1, n2 = rng.integers(1, 12, 2)
        a, b = rnd(rng, n1), rnd(rng, n2)
        eps = rng.random() * 0.6
        thr = O.eps_threshold(eps)
        assert _edr_short_rule(a, b, thr) == O.count("edr", a, b, eps)


def _int_forms(a, b, thr):
    """EDR and LCSS counts with the FP32 kernel's integer updates:
    N = max(N_diag + 1 + m, N_up, N_left) on N = (i + j) - E with the
    boundaries N[0][j] = j, N[i][0] = i (the count is n1 + n2 - N), and
    L = max(L_up, L_left, L_diag + m)."""
    n1, n2 = len(a), len(b)
    N = np.zeros((n1 + 1, n2 + 1), dtype=np.int64)
    N[0, :] = np.arange(n2 + 1)
    N[:, 0] = np.arange(n1 + 1)
    L = np.zeros((n1 + 1, n2 + 1), dtype=np.int64)
    for i in range(1, n1 + 1):
        for j in range(1, n2 + 1):
            d2 = (a[i - 1, 0] - b[j - 1, 0]) ** 2 + (a[i - 1, 1] - b[j - 1, 1]) ** 2
            m = int(d2 < thr)
            N[i, j] = max(N[i - 1, j - 1] + 1 + m, N[i - 1, j], N[i, j - 1])
            L[i, j] = max(L[i - 1, j - 1] + m, L[i - 1, j], L[i, j - 1])
    return n1 + n2 - int(N[n1, n2]), int(L[n1, n2])


def test_int_update_forms(rng):
    """The FP32 kernel's integer EDR (shifted) and LCSS (max3) updates give
    the oracle's counts, at match densities from none to all."""
    for _ in range(300):
        n1, n2 = rng.integers(1, 12, 2)
        a, b = rnd(rng, n1), rnd(rng, n2)
        eps = rng.random() * 1.2
        thr = O.eps_threshold(eps)
        e, l = _int_forms(a, b, thr)
        assert e == O.count("edr", a, b, eps)
        assert l == O.count("lcss", a, b, eps)


def test_fp32_bit_pattern_match_and_band():
    """For non-negative FP32 d2 and 0 <= lo <= hi: the sign bit of
    bits(d2) - bits(lo) is d2 < lo, and bits(d2) - bits(lo) < bits(hi) -
    bits(lo) (unsigned) is lo <= d2 < hi (zero, subnormals, the band's
    edges, huge values and infinity included)."""
    f32 = np.float32
    rng = np.random.default_rng(5)
    d2 = np.concatenate([
        np.array([0.0, 1e-45, 1e-40, 1e-38, 6.25e-6, 1.0, 3.4e38, np.inf], dtype=f32),
        rng.random(20000).astype(f32) * f32(2e-5),
        (rng.random(2000) * 1e30).astype(f32),
    ])
    bands = [(0.0, 0.0), (np.inf, np.inf), (6.2e-6, 6.3e-6), (1e-45, 1e-40), (1.0, 1.0), (0.0, 3.4e38)]
    for lo, hi in bands:
        lo, hi = f32(lo), f32(hi)
        # the band's own edges and their neighbours
        edges = np.array([lo, hi, np.nextafter(lo, f32(0)), np.nextafter(lo, f32(np.inf)),
                          np.nextafter(hi, f32(0)), np.nextafter(hi, f32(np.inf))], dtype=f32)
        x = np.concatenate([d2, edges])
        xb = x.view(np.uint32)
        lob, hib = np.array([lo], dtype=f32).view(np.uint32)[0], np.array([hi], dtype=f32).view(np.uint32)[0]
        t = (xb - lob).astype(np.uint32)
        assert np.array_equal((t >> 31) == 1, x < lo), (lo, hi)
        assert np.array_equal(t < np.uint32(hib - lob), (x >= lo) & (x < hi)), (lo, hi)


@pytest.mark.parametrize("n1,n2", [(1, 1), (1, 3), (3, 2), (3, 3), (4, 3)])
def test_st_bruteforce(rng, n1, n2):
    """STEDR / STLCSS (core/stedr.py:44-46, core/stlcss.py:41-42): a cell
    matches when the points are within eps AND their times within delta."""
    for eps, delta in ((0.3, 0.5), (0.6, 0.2), (0.6, 10.0)):
        thr = O.eps_threshold(eps)
        for _ in range(4):
            a, b = rnd(rng, n1), rnd(rng, n2)
            ta, tb = rng.random(n1), rng.random(n2)
            A = [(p[0], p[1], t) for p, t in zip(a, ta)]
            B = [(p[0], p[1], t) for p, t in zip(b, tb)]
            match = lambda p, q: (p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 < thr and abs(p[2] - q[2]) < delta
            assert O.st_count("stedr", a, ta, b, tb, eps, delta) == BF.edr_count(A, B, match)
            assert O.st_count("stlcss", a, ta, b, tb, eps, delta) == BF.lcss_count(A, B, match)


def test_st_limits(rng):
    """delta = inf reduces STEDR/STLCSS to EDR/LCSS; delta <= 0 leaves no
    match (STEDR count min(n1, n2) under the zero boundary, STLCSS 0)."""
    for _ in range(10):
        n1, n2 = rng.integers(1, 30, 2)
        a, b = rnd(rng, n1), rnd(rng, n2)
        ta, tb = rng.random(n1) * 100, rng.random(n2) * 100
        for eps in (0.05, 0.2):
            assert O.st_count("stedr", a, ta, b, tb, eps, np.inf) == O.count("edr", a, b, eps)
            assert O.st_count("stlcss", a, ta, b, tb, eps, np.inf) == O.count("lcss", a, b, eps)
            assert O.st_count("stedr", a, ta, b, tb, eps, 0.0) == min(n1, n2)
            assert O.st_count("stlcss", a, ta, b, tb, eps, 0.0) == 0
        # identical points, times 10 apart shifted by exactly delta: the strict
        # test never matches; shifted by less, exactly the diagonal does
        ti = 10.0 * np.arange(n1)
        assert O.st_count("stlcss", a, ti, a, ti + 5.0, 1.0, 5.0) == 0
        assert O.st_count("stlcss", a, ti, a, ti + 4.0, 1e-9, 5.0) == n1


def test_st_pairs_matches_single(rng):
    from trajsimimeasures_amd.data import lengths_uniform, random_walks
    ts = random_walks(rng, lengths_uniform(rng, 40, 1, 30))
    t = np.concatenate([np.cumsum(rng.random(n)) * 10 for n in ts.lengths])
    pa, pb = rng.integers(0, 40, 60), rng.integers(0, 40, 60)
    for m in ("stedr", "stlcss"):
        vals, cnts = O.pairs_st(m, ts.xy, t, ts.offsets, pa, pb, 0.02, 15.0)
        for k in range(60):
            a, b = ts.traj(pa[k]), ts.traj(pb[k])
            ta = t[ts.offsets[pa[k]]:ts.offsets[pa[k] + 1]]
            tb = t[ts.offsets[pb[k]]:ts.offsets[pb[k] + 1]]
            c = O.st_count(m, a, ta, b, tb, 0.02, 15.0)
            assert cnts[k] == c
            want = c / max(len(a), len(b)) if m == "stedr" else 1.0 - c / min(len(a), len(b))
            assert vals[k] == want


def _within_time(e1p1, e1p2, e2p1, e2p2, eps, n=200001):
    """Measure of {t : |x1(t) - x2(t)| <= eps} over the common time span, by
    dense sampling of the two linear motions."""
    t0, t1 = max(e1p1[2], e2p1[2]), min(e1p2[2], e2p2[2])
    if t1 <= t0:
        return 0.0
    t = np.linspace(t0, t1, n)

    def pos(p, q):
        w = (t - p[2]) / (q[2] - p[2])
        return p[0] + w * (q[0] - p[0]), p[1] + w * (q[1] - p[1])
    x1, y1 = pos(e1p1, e1p2)
    x2, y2 = pos(e2p1, e2p2)
    inside = (x1 - x2) ** 2 + (y1 - y2) ** 2 <= eps * eps
    return inside.mean() * (t1 - t0)


def test_cdd_matches_sampled_duration(rng):
    """core/cdds.py:21-73: the closed-form duration within eps equals a
    dense sampling of the two motions (moving and parallel cases)."""
    for _ in range(40):
        ts = np.sort(rng.random(4) * 10)
        e1p1, e1p2 = [*rng.random(2), ts[0]], [*rng.random(2), ts[2]]
        e2p1, e2p2 = [*rng.random(2), ts[1]], [*rng.random(2), ts[3]]
        for eps in (0.2, 0.5, 1.0):
            want = _within_time(e1p1, e1p2, e2p1, e2p2, eps)
            span = max(0.0, min(e1p2[2], e2p2[2]) - max(e1p1[2], e2p1[2]))
            assert abs(O.cdd(e1p1, e1p2, e2p1, e2p2, eps) - want) <= 1e-4 * span + 1e-12
    # parallel motion at constant distance 0.5: all or nothing
    p, q = [0.0, 0.0, 0.0], [1.0, 0.0, 2.0]
    r, s = [0.0, 0.5, 1.0], [1.0, 0.5, 3.0]
    assert O.cdd(p, q, [0.0, 0.5, 0.0], [1.0, 0.5, 2.0], 0.6) == 2.0
    assert O.cdd(p, q, [0.0, 0.5, 0.0], [1.0, 0.5, 2.0], 0.4) == 0.0
    # disjoint or touching time spans, and zero-duration segments, give 0
    assert O.cdd(p, q, [0, 0, 2.0], [1, 0, 3.0], 1.0) == 0.0
    assert O.cdd(p, [1.0, 0.0, 0.0], r, s, 1.0) == 0.0


def test_cdds_closed_forms_and_sweep(rng):
    """A trajectory against itself is within any eps > 0 for its whole
    duration; and with increasing timestamps the all-segment-pairs sum of
    core/cdds.py equals the merge sweep of core_cpu/cdds.pyx:80-104."""
    for _ in range(10):
        n1, n2 = rng.integers(2, 25, 2)
        a, b = rnd(rng, n1), rnd(rng, n2)
        ta = np.cumsum(rng.random(n1) + 0.1)
        tb = np.cumsum(rng.random(n2) + 0.1) + rng.random()
        assert O.cdds_single(a, ta, a, ta, 1e-9) == pytest.approx(ta[-1] - ta[0], rel=1e-12)
        for eps in (0.1, 0.4):
            i = j = 0
            sweep = 0.0
            while i < n1 - 1 and j < n2 - 1:
                sweep += O.cdd([*a[i], ta[i]], [*a[i + 1], ta[i + 1]], [*b[j], tb[j]], [*b[j + 1], tb[j + 1]], eps)
                if ta[i + 1] < tb[j + 1]:
                    i += 1
                elif tb[j + 1] < ta[i + 1]:
                    j += 1
                else:
                    i += 1
                    j += 1
            assert O.cdds_single(a, ta, b, tb, eps) == pytest.approx(sweep, rel=1e-12, abs=1e-15)
            assert O.cdds_single(b, tb, a, ta, eps) == pytest.approx(sweep, rel=1e-9, abs=1e-12)
