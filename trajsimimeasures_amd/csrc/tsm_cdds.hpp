// tsm_cdds.hpp -- CDDS (core/cdds.py:11-124): for two trajectories of
// timestamped points, the total time their linearly moving positions spend
// within eps of each other, summed over every pair of segments.
//
// Per segment, x(t) = xk t + xb and y(t) = yk t + yb (core/cdds.py:11-18),
// computed where a segment pair is evaluated (cdds_line).  One wavefront per
// pair: lanes take 64 segments of the first trajectory at a time and add the
// time their segment stays within eps of each segment of the second
// (core/cdds.py:21-73), in segment order.  When the second trajectory's
// timestamps are non-decreasing (checked per pair), a lane visits only the
// segments that overlap its own in time: a binary search finds the first one
// ending after its start, and it stops at the first one starting at or after
// its end.  Every skipped segment would add exactly +0 (no time overlap), so
// the sums are the same numbers.  Otherwise the second trajectory's segments
// are walked wave-uniformly, all of them.  The per-segment sums are then
// added in segment order by one lane, the oracle's order, so results match it
// bit for bit.  A segment without duration contributes 0, as in
// core_cpu/cdds.pyx:26 (the reference GPU kernel would divide by zero).
// Everything is FP64.
#pragma once

#include "tsm_device.hpp"

namespace tsm {

struct CArgs {
    const double *xy;
    const double *t;
    const int64_t *offs;
    PairSpace ps;
    int64_t npairs;
    int64_t nset;  // trajectories in the set
    double eps;
    double *out;
};

// Segment (i, i + 1) as slope / intercept in time, x(t) = xk t + xb and
// y(t) = yk t + yb (core/cdds.py:11-18), returned as (xk, xb, yk, yb); a
// zero-duration segment gets zeros, which the time test never reads (it
// rejects by the timestamps first).  Computed where a segment pair is
// evaluated: a pass writing these for every point would cost more HBM
// traffic than the measure itself.
__device__ __forceinline__ double4 cdds_line(const double *xy, const double *t, int64_t i) {
    double4 l = make_double4(0.0, 0.0, 0.0, 0.0);
    if (t[i + 1] != t[i]) {
        const double x1 = xy[2 * i], y1 = xy[2 * i + 1], t1 = t[i];
        const double x2 = xy[2 * i + 2], y2 = xy[2 * i + 3], t2 = t[i + 1];
        const double dt = t2 - t1;
        l.z = (y2 - y1) / dt;
        l.w = (y1 * t2 - y2 * t1) / dt;
        l.x = (x2 - x1) / dt;
        l.y = (x1 * t2 - x2 * t1) / dt;
    }
    return l;
}

// The oracle's cdd (core/cdds.py:21-73) on segments as cdds_line gives them.
__device__ __forceinline__ double cdd_line(double4 A, double as, double ae, double4 B, double bs, double be,
                                           double se) {
    if (ae <= bs || be <= as || as == ae || bs == be) return 0.0;
    const double dxk = A.x - B.x, dyk = A.z - B.z, dxb = A.y - B.y, dyb = A.w - B.w;
    const double a = dxk * dxk + dyk * dyk;
    const double b = 2.0 * (dxk * dxb + dyk * dyb);
    double c = dxb * dxb + dyb * dyb;
    const double t_start = as > bs ? as : bs;
    const double t_end = ae < be ? ae : be;
    if (a == 0.0 && b == 0.0 && c == 0.0) return t_end - t_start;
    if (a == 0.0 && b == 0.0) return c <= se ? t_end - t_start : 0.0;
    if (a == 0.0) {
        const double t1 = -1.0 * c / b, t2 = (se - c) / b;
        if (t1 < t2) {
            const double lo = t1 > t_start ? t1 : t_start, hi = t2 < t_end ? t2 : t_end;
            return hi > lo ? hi - lo : 0.0;
        }
        if (t2 < t1) {
            const double lo = t2 > t_start ? t2 : t_start, hi = t1 < t_end ? t1 : t_end;
            return hi > lo ? hi - lo : 0.0;
        }
        return 0.0;
    }
    c -= se;
    const double theta = b * b - 4.0 * a * c;
    if (theta > 0.0) {
        const double ts = sqrt(theta);
        const double t1 = (-1.0 * ts - b) / 2.0 / a, t2 = (ts - b) / 2.0 / a;
        const double lo = t1 > t_start ? t1 : t_start, hi = t2 < t_end ? t2 : t_end;
        return hi > lo ? hi - lo : 0.0;
    }
    return 0.0;
}

__device__ __forceinline__ double uni_d(double v) {
    const int64_t b = __builtin_bit_cast(int64_t, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double lane_d(double v, int l) {
    const int64_t b = __builtin_bit_cast(int64_t, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

__global__ __launch_bounds__(kWave) void cdds_kernel(CArgs a) {
    const int lane = threadIdx.x;
    const double se = a.eps * a.eps;
    for (int64_t p = blockIdx.x; p < a.npairs; p += gridDim.x) {
        int64_t ia, ib;
        decode_pair(a.ps, p, ia, ib);
        if (ia < 0 || ia >= a.nset || ib < 0 || ib >= a.nset) {  // index outside the set: NaN, no access
            if (lane == 0) a.out[p] = __builtin_bit_cast(double, 0x7ff8000000000000ull);
            continue;
        }
        const int64_t oa = a.offs[ia], ob = a.offs[ib];
        const int na = (int)(a.offs[ia + 1] - oa), nb = (int)(a.offs[ib + 1] - ob);
        const double *tb = a.t + ob;
        bool mono = true;
        for (int j = lane; j + 1 < nb; j += kWave) mono = mono && tb[j + 1] >= tb[j];
        const bool window = __builtin_amdgcn_ballot_w64(!mono) == 0;
        double total = 0.0;  // wave-uniform
        for (int base = 0; base + 1 < na; base += kWave) {
            const int i = base + lane;
            const bool ok = i + 1 < na;
            const int ic = ok ? i : 0;
            const double4 A = cdds_line(a.xy + 2 * oa, a.t + oa, ic);
            const double as = a.t[oa + ic], ae = a.t[oa + ic + (ok ? 1 : 0)];
            double d = 0.0;
            if (window) {
                // first segment of the second trajectory that ends after as (nb - 1: none)
                int lo = 0, hi = nb - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (tb[mid + 1] > as) hi = mid;
                    else lo = mid + 1;
                }
                for (int j = lo; j + 1 < nb && tb[j] < ae; ++j)
                    d += cdd_line(A, as, ae, cdds_line(a.xy + 2 * ob, tb, j), tb[j], tb[j + 1], se);
            } else {
                for (int j = 0; j + 1 < nb; ++j) {
                    const double4 Bv = cdds_line(a.xy + 2 * ob, tb, j);
                    const double4 B = make_double4(uni_d(Bv.x), uni_d(Bv.y), uni_d(Bv.z), uni_d(Bv.w));
                    const double bs = uni_d(a.t[ob + j]), be = uni_d(a.t[ob + j + 1]);
                    d += cdd_line(A, as, ae, B, bs, be, se);
                }
            }
            // per-segment sums in segment order, as the oracle adds them
            const int cnt = na - 1 - base < kWave ? na - 1 - base : kWave;
            for (int l = 0; l < cnt; ++l) total += lane_d(d, l);
        }
        if (lane == 0) a.out[p] = total;
    }
}

}  // namespace tsm
