"""Pins the SAR oracle (oracle/tsm_oracle.c, tsmo_sar), the reference GPU
form of core/sar.py:10-212, from several sides:

- a list-based restatement written the way core/sar.py and
  core_cpu/sar.pyx hold the data (sample lists, float32 storage, slices,
  Python's // and %), compared value for value on random trajectories with
  increasing and with scrambled timestamps;
- closed forms: identical trajectories, a constant spatial offset, a
  stationary object seen with a time shift, linear motion sampled at
  different times, no temporal overlap;
- hand cases for the bins, whose ranges overlap by one sample in the GPU form
  (core/sar.py:195) but not in the CPU form (core_cpu/sar.pyx:125-130);
- the relation to the CPU form: equal when threads_per_traj is 1 and no
  binning happens.

No reference code is imported or run."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _align(pts, ts):
    """Samples of pts [n, 3] at multiples of ts (core/sar.py:31-59)."""
    out = []
    for (x1, y1, t1), (x2, y2, t2) in zip(pts[:-1], pts[1:]):
        s, e = math.ceil(t1 / ts), math.ceil(t2 / ts)
        if s < e:
            xk = (x2 - x1) / (t2 - t1)
            xb = (x1 * t2 - x2 * t1) / (t2 - t1)
            yk = (y2 - y1) / (t2 - t1)
            yb = (y1 * t2 - y2 * t1) / (t2 - t1)
            for tt in range(s * ts, e * ts, ts):
                out.append((np.float32(xk * tt + xb), np.float32(yk * tt + yb), np.float32(tt)))
    if pts[-1][2] % ts == 0:
        out.append(tuple(np.float32(v) for v in pts[-1]))
    return out


def sar_lists(p, q, deps, ts, M, R):
    """The GPU form of SAR over Python lists (see module doc)."""
    p, q = [tuple(r) for r in p], [tuple(r) for r in q]
    span = None
    i = j = 0
    while i < len(p) - 1 and j < len(q) - 1:
        (a0, a1), (b0, b1) = (p[i][2], p[i + 1][2]), (q[j][2], q[j + 1][2])
        if not (a1 <= b0 or b1 <= a0 or a0 == a1 or b0 == b1):
            span = [i, i + 1, j, j + 1] if span is None else [span[0], i + 1, span[2], j + 1]
        if a1 < b1:
            i += 1
        elif b1 < a1:
            j += 1
        else:
            i, j = i + 1, j + 1
    if span is None:
        return 0.0
    A = _align(p[span[0]:span[1] + 1], ts)
    B = _align(q[span[2]:span[3] + 1], ts)
    if not A or not B:
        return 0.0
    lo1, hi1, lo2, hi2 = 0, len(A), 0, len(B)
    ae, be = A[-1][2], B[-1][2]
    if ae < be:
        hi2 -= min(int((be - ae) // ts), hi2 - lo2)
    elif ae > be:
        hi1 -= min(int((ae - be) // ts), hi1 - lo1)
    if hi1 - lo1 == 0 or hi2 - lo2 == 0:
        return 0.0
    as_, bs = A[0][2], B[0][2]
    if as_ < bs:
        lo1 += min(int((bs - as_) // ts), len(A))
    elif as_ > bs:
        lo2 += min(int((as_ - bs) // ts), len(B))
    A, B = A[lo1:hi1], B[lo2:hi2]
    sz = min(len(A), len(B))
    if sz <= 0:
        return 0.0
    if M >= sz:
        hits = sum(1 for u, v in zip(A[:sz], B[:sz])
                   if np.sqrt((v[0] - u[0]) * (v[0] - u[0]) + (v[1] - u[1]) * (v[1] - u[1])) <= deps)
        return float(R * hits * ts)
    k, m = divmod(sz, M)
    total = 0
    for b in range(M):
        st = b * k + min(b, m)
        en = min((b + 1) * k + min(b + 1, m) + 1, sz)
        mx = [0.0, 0.0, 0.0, 0.0]
        for u, v in zip(A[st:en], B[st:en]):
            mx = [mx[0] + float(u[0]), mx[1] + float(u[1]), mx[2] + float(v[0]), mx[3] + float(v[1])]
        mx = [c / (en - st) for c in mx]
        if math.sqrt((mx[2] - mx[0]) ** 2 + (mx[3] - mx[1]) ** 2) <= deps:
            total += (en - st) * ts
    return float(total)


def _traj(rng, n, t0, interval, jitter, scramble=False):
    xy = np.cumsum(rng.normal(0, 0.01, (n, 2)), axis=0) + rng.random(2)
    gaps = interval + jitter * (2 * rng.random(n) - 1)
    gaps[0] = 0
    t = t0 + np.cumsum(gaps)
    if rng.random() < 0.3:
        t = np.round(t)  # some timestamps on whole seconds (and grid multiples)
    if scramble:
        k = rng.integers(1, 4)
        for _ in range(k):
            a, b = rng.integers(0, n, 2)
            t[a], t[b] = t[b], t[a]
    return np.column_stack([xy, t])


@pytest.mark.parametrize("scramble", [False, True])
def test_oracle_matches_list_restatement(scramble):
    rng = np.random.default_rng(11 if scramble else 7)
    for case in range(400):
        n1, n2 = rng.integers(1, 40, 2)
        p = _traj(rng, n1, rng.random() * 200, 15, 5, scramble)
        q = _traj(rng, n2, rng.random() * 200, 12, 8, scramble)
        ts = int(rng.choice([1, 7, 10, 30]))
        M = int(rng.choice([1, 3, 10, 50]))
        R = int(rng.choice([1, 64]))
        deps = float(rng.choice([0.005, 0.02, 0.1]))
        want = sar_lists(p, q, deps, ts, M, R)
        got = O.sar_single(p[:, :2], p[:, 2], q[:, :2], q[:, 2], deps, ts, M, R)
        assert got == want, (case, got, want)


def test_identity_and_offset_closed_forms():
    rng = np.random.default_rng(3)
    for n in (2, 5, 17, 64):
        xy = rng.random((n, 2))
        t = 10.0 * np.arange(n)
        # identical: every one of the n grid samples matches
        assert O.sar_single(xy, t, xy, t, 1e-9, 10, n, 64) == 64 * n * 10
        assert O.sar_single(xy, t, xy, t, 1e-9, 10, n, 1) == n * 10
        # constant offset 0.5 within / beyond distance_eps, bins (M < n):
        # bins cover the n samples plus M - 1 shared ones
        for M in range(1, n):
            assert O.sar_single(xy, t, xy + [0.5, 0.0], t, 1.0, 10, M, 64) == (n + M - 1) * 10
            assert O.sar_single(xy, t, xy + [0.5, 0.0], t, 0.25, 10, M, 64) == 0.0


def test_stationary_time_shift():
    """A stationary object seen over [0, 10(n-1)] and over the same span
    shifted by s grid steps: n - s common samples, all within eps."""
    n = 30
    xy = np.full((n, 2), 0.25)
    t = 10.0 * np.arange(n)
    for s in range(0, n):
        want = 0.0 if s == n - 1 else 3 * (n - s) * 10.0
        got = O.sar_single(xy, t, xy, t + 10 * s, 1e-6, 10, 1000, 3)
        if s == n - 1:  # the spans touch at one instant: no open overlap
            assert got == 0.0
        else:
            assert got == want, (s, got, want)


def test_linear_motion_resampled():
    """Motion x = 0.001 t sampled at different instants: every common grid
    time matches; count = grid times in the common span."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        ta = np.sort(rng.uniform(0, 600, rng.integers(3, 30)))
        tb = np.sort(rng.uniform(0, 600, rng.integers(3, 30)))
        a = np.column_stack([0.001 * ta, np.zeros_like(ta)])
        b = np.column_stack([0.001 * tb, np.zeros_like(tb)])
        got = O.sar_single(a, ta, b, tb, 1e-5, 10, 10 ** 6, 1)
        lo, hi = max(ta[0], tb[0]), min(ta[-1], tb[-1])
        if lo >= hi:
            assert got == 0.0
            continue
        # samples of a: grid times in [ceil(ta0/10)*10, ta_end) (ta_end is not
        # on the grid here), likewise b; common ones
        ga = set(range(math.ceil(ta[0] / 10), math.ceil(ta[-1] / 10)))
        gb = set(range(math.ceil(tb[0] / 10), math.ceil(tb[-1] / 10)))
        assert got == 10.0 * len(ga & gb)


def test_bins_hand_case():
    """5 samples, t = 0..40; the second trajectory moves to x = 5 for its
    last two.  M = 2 gives bins [0, 4) and [3, 5) (div 2, mod 1, "+ 1"):
    bin 0 mean offset 5/4, bin 1 offset 5."""
    t = 10.0 * np.arange(5)
    a = np.zeros((5, 2))
    b = np.zeros((5, 2))
    b[3:, 0] = 5.0
    assert O.sar_single(a, t, b, t, 1.0, 10, 2, 64) == 0.0
    assert O.sar_single(a, t, b, t, 2.0, 10, 2, 64) == 40.0
    assert O.sar_single(a, t, b, t, 6.0, 10, 2, 64) == 60.0
    # M >= 5: per-sample count 3, times threads_per_traj
    assert O.sar_single(a, t, b, t, 1.0, 10, 5, 64) == 64 * 3 * 10
    assert O.sar_single(a, t, b, t, 1.0, 10, 5, 1) == 3 * 10
    # M = 1 with 5 samples: one bin over all five (clipped at the end)
    assert O.sar_single(a, t, b, t, 2.0, 10, 1, 64) == 50.0
    assert O.sar_single(a, t, b, t, 1.9, 10, 1, 64) == 0.0


def test_no_overlap_and_degenerate():
    a = np.zeros((4, 2))
    t = np.array([0.0, 10, 20, 30])
    assert O.sar_single(a, t, a, t + 30, 1.0, 10, 10, 64) == 0.0   # touch at t = 30 only
    assert O.sar_single(a, t, a, t + 100, 1.0, 10, 10, 64) == 0.0
    assert O.sar_single(a[:1], t[:1], a, t, 1.0, 10, 10, 64) == 0.0  # one point: no segment
    assert O.sar_single(a, np.zeros(4), a, t, 1.0, 10, 10, 64) == 0.0  # no duration
    # overlap inside one grid gap: no sample on one side
    assert O.sar_single(a[:2], np.array([1.0, 9.0]), a[:2], np.array([2.0, 8.0]), 1.0, 10, 10, 64) == 0.0


def test_trim_uses_floor_division_of_float32_differences():
    """Ends 3 grid steps apart are trimmed by exactly 3 samples; with times
    off the grid the stretches hold 4 samples each; at 2^24 s and beyond the
    float32 sample times are rounded, which the trim then sees."""
    for base in (0.0, 1000.0, 2.0 ** 23, 2.0 ** 24 + 2):
        t = base + 10.0 * np.arange(8)
        a = np.zeros((8, 2))
        got = O.sar_single(a, t, a[:5], t[:5], 1.0, 10, 100, 1)
        assert got == sar_lists(np.column_stack([a, t]), np.column_stack([a[:5], t[:5]]), 1.0, 10, 100, 1)
        if base < 2.0 ** 24:
            assert got == (5 if base % 10 == 0 else 4) * 10.0


def test_cpu_form_relation():
    """With threads_per_traj = 1 and no binning (M >= samples) the GPU form
    equals core_cpu/sar.pyx's count * time_eps: checked here on grid-aligned
    data whose samples are the points themselves."""
    rng = np.random.default_rng(9)
    for _ in range(40):
        n = int(rng.integers(2, 30))
        t = 10.0 * np.arange(n)
        a = np.round(rng.random((n, 2)) * 64) / 64  # float32-exact coordinates
        b = np.round(rng.random((n, 2)) * 64) / 64
        d = np.sqrt(((a - b) ** 2).sum(1))
        u = np.unique(d)
        if u.size < 2:
            continue
        k = u.size // 2  # eps halfway between two distances: no rounding can cross it
        eps = float(0.5 * (u[k - 1] + u[k]))
        cpu = int(np.sum(d <= eps)) * 10
        assert O.sar_single(a, t, b, t, eps, 10, n, 1) == cpu
