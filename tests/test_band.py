"""The guard band of the FP32 match test of EDR/LCSS (count_band in
csrc/tsm_host.hpp): with FP32-rounded coordinates, dx' = fl(ax' - bx'),
d2' = fma(dy', dy', fl(dx' * dx')), the FP32 squared distance differs from
the FP64 one by less than rel * thr for points near the threshold, where
  rel = 7u + 12u U / sqrt(thr) + 20u^2 (U + sqrt(thr))^2 / thr,  u = 2^-24.
Checked here on adversarial samples (coordinates up to U, squared distances
within a few percent of thr), so a dropped term in the bound fails."""
import numpy as np

U32 = 2.0 ** -24


def rel_bound(thr, U):
    e = np.sqrt(thr)
    return 7 * U32 + 12 * U32 * U / e + 20 * U32 * U32 * (U + e) ** 2 / thr


def fp32_d2(ax, ay, bx, by):
    f = np.float32
    dx = (f(ax) - f(bx)).astype(np.float32)
    dy = (f(ay) - f(by)).astype(np.float32)
    p = (dx * dx).astype(np.float32)
    # fma: dy*dy is exact in float64 (24 x 24 bits), one rounding to float32
    return (dy.astype(np.float64) ** 2 + p.astype(np.float64)).astype(np.float32).astype(np.float64)


def test_band_bounds_fp32_error():
    rng = np.random.default_rng(0)
    worst = 0.0
    for U in (1.0, 5.0, 100.0):
        for eps in (0.0025, 0.01, 0.5):
            thr = eps * eps
            n = 200_000
            ax = rng.uniform(-U, U, n)
            ay = rng.uniform(-U, U, n)
            th = rng.uniform(0, 2 * np.pi, n)
            r = eps * rng.uniform(0.97, 1.03, n)
            bx = np.clip(ax + r * np.cos(th), -U, U)
            by = np.clip(ay + r * np.sin(th), -U, U)
            d64 = (ax - bx) * (ax - bx) + (ay - by) * (ay - by)
            d32 = fp32_d2(ax, ay, bx, by)
            err = np.abs(d32 - d64).max() / thr
            bound = rel_bound(thr, U)
            assert err < bound, (U, eps, err, bound)
            worst = max(worst, err / bound)
    # the bound is not vacuous: the worst sample uses a good part of it
    assert worst > 0.05


def test_band_decisions_match_fp64():
    """Decisions outside [thr (1 - 3 rel), thr (1 + 3 rel)) agree with FP64."""
    rng = np.random.default_rng(1)
    U, eps = 5.0, 0.0025
    thr = eps * eps
    B = 3 * rel_bound(thr, U)
    n = 400_000
    ax = rng.uniform(0, U, n)
    ay = rng.uniform(0, U, n)
    th = rng.uniform(0, 2 * np.pi, n)
    r = eps * rng.uniform(0.99, 1.01, n)
    bx, by = ax + r * np.cos(th), ay + r * np.sin(th)
    d64 = (ax - bx) * (ax - bx) + (ay - by) * (ay - by)
    d32 = fp32_d2(ax, ay, bx, by)
    lo, hi = thr * (1 - B), thr * (1 + B)
    sure_match = d32 < lo
    sure_miss = d32 >= hi
    assert np.all(d64[sure_match] < thr) and np.all(d64[sure_miss] >= thr)
    # and the band is narrow: most samples are decided in FP32
    assert (sure_match | sure_miss).mean() > 0.2


def test_band_margin_report():
    rng = np.random.default_rng(2)
    U, eps = 5.0, 0.0025
    thr = eps * eps
    ax = rng.uniform(-U, U, 100000)
    bx = ax + eps
    d32 = fp32_d2(ax, 0.0 * ax, bx, 0.0 * bx)
    d64 = (ax - bx) ** 2
    assert (np.abs(d32 - d64) / thr).max() < rel_bound(thr, U)
