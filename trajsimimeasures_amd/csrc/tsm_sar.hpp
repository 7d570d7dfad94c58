// tsm_sar.hpp -- SAR (core/sar.py:10-286): how long two timestamped
// trajectories stay within distance_eps of each other, compared on a common
// time grid (every time_eps seconds) over the stretch where both exist.
//
// The reference kernel, per pair (core/sar.py:62-212):
//  1. a merge of both trajectories' segments in order of end time finds the
//     first and last segments that overlap the other trajectory in time;
//  2. each overlapping stretch is sampled at every multiple of time_eps
//     (linear interpolation on the segment holding it, plus the stretch's
//     last point when it lies on the grid), the samples stored in FP32;
//  3. the later end and then the earlier start lose (difference // time_eps)
//     samples, so both stretches cover the same grid times;
//  4. with sz samples left and target_length M >= sz, count the sample pairs
//     within distance_eps; every one of the reference's threads_per_traj
//     threads stores that same count and the host sums them (core/sar.py:
//     179-181, 236, 285), so the result is threads_per_traj * count *
//     time_eps.  With M < sz, M bins of div or div + 1 samples (div, mod =
//     sz // M, sz % M) plus one: bin b is [b*div + min(b, mod),
//     (b+1)*div + min(b+1, mod) + 1) (core/sar.py:194-195), and a bin whose
//     FP64 mean positions are within distance_eps adds its sample count times
//     time_eps.
// Where the reference reads outside the samples it holds, the result is
// defined here: the last bin ends at sz (its "+ 1" sample lies past the
// stretch), and stretches of different lengths (timestamps past 2^24 s,
// whose float32 copies round, or decreasing ones) use the shorter length.
// No overlap gives 0 (the reference leaves the slot unwritten); a stretch of
// more than 2^24 samples gives NaN.  Timestamps must not be NaN.
//
// MI355X mapping: one wavefront per pair.  When both trajectories'
// timestamps are non-decreasing (the normal case), everything is data
// parallel: the overlap bounds are the first / last segments that overlap
// the other trajectory's time span (the merge visits every overlapping
// segment pair of sorted sequences, so these are the merge's bounds); the
// samples of a stretch are the grid indices ceil(t_is / time_eps) ..
// ceil(t_ie / time_eps) - 1 in order, and sample m sits on the segment found
// by a binary search over ceil(t_k / time_eps); lanes compute 64 samples of
// both trajectories at a time.  Counts reduce with a ballot; bin means are
// summed in sample order (the reference's FP64 order) by the lane that owns
// the bin, reading the chunk's samples from LDS.  Any decreasing timestamp
// sends the pair to a sequential path on lane 0 that follows the steps above
// literally.
#pragma once

#include "tsm_device.hpp"

namespace tsm {

struct SArgs {
    const double *xy;
    const double *t;
    const int64_t *offs;
    PairSpace ps;
    int64_t npairs;
    int64_t nset;
    double deps;  // distance_eps
    int64_t ts;   // time_eps, > 0
    int64_t M;    // target_length, >= 1
    int64_t R;    // threads_per_traj of the reference (replicates the per-sample count)
    double *out;
};

constexpr int64_t kSarMaxSamples = int64_t(1) << 24;
constexpr int kSarStage = 1024;  // segments of a stretch whose grid indices are staged in LDS

__device__ __forceinline__ double sar_nan() { return __builtin_bit_cast(double, 0x7ff8000000000000ull); }

// float -> int64 as the reference's math.ceil / int() results: truncation,
// saturating
__device__ __forceinline__ int64_t sar_i64(double v) {
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

__device__ __forceinline__ int64_t sar_grid(double t, double tsd) { return sar_i64(ceil(t / tsd)); }

// Python's float floor division a // b (fmod first, so that an exact
// multiple never rounds down), as Numba evaluates core/sar.py:129-150.
__device__ __forceinline__ double sar_floordiv(double a, double b) {
    const double mod = fmod(a, b);
    double div = (a - mod) / b;
    if (mod != 0.0 && ((b < 0.0) != (mod < 0.0))) div -= 1.0;
    if (div == 0.0) return copysign(0.0, a / b);
    double fl = floor(div);
    if (div - fl > 0.5) fl += 1.0;
    return fl;
}

// Interpolated position on segment (k, k + 1) at time tt, the expressions of
// core/sar.py:15-25, rounded to FP32 like the reference's sample buffer.
__device__ __forceinline__ float2 sar_interp(const double *xy, const double *t, int64_t k, int64_t tt) {
    const double t1 = t[k], t2 = t[k + 1];
    const double x1 = xy[2 * k], y1 = xy[2 * k + 1], x2 = xy[2 * k + 2], y2 = xy[2 * k + 3];
    const double yk = (y2 - y1) / (t2 - t1), yb = (y1 * t2 - y2 * t1) / (t2 - t1);
    const double xk = (x2 - x1) / (t2 - t1), xb = (x1 * t2 - x2 * t1) / (t2 - t1);
    const double td = (double)tt;
    return make_float2((float)(xk * td + xb), (float)(yk * td + yb));
}

// Trimmed sample ranges of both stretches (core/sar.py:118-156) from their
// sample counts and first / last sample times.  Returns sz (<= 0: nothing
// to compare) and the start indices.
__device__ __forceinline__ int64_t sar_trim(int64_t c1, int64_t c2, float as, float ae, float bs, float be,
                                            double tsd, int64_t &s1, int64_t &s2) {
    int64_t e1 = c1, e2 = c2, z1 = c1, z2 = c2;
    s1 = 0;
    s2 = 0;
    if (ae == be) {
    } else if (ae < be) {
        const int64_t k = sar_i64(sar_floordiv((double)(float)(be - ae), tsd));
        e2 -= k < z2 ? k : z2;
        z2 = e2 - s2;
    } else if (ae > be) {
        const int64_t k = sar_i64(sar_floordiv((double)(float)(ae - be), tsd));
        e1 -= k < z1 ? k : z1;
        z1 = e1 - s1;
    }
    if (z1 == 0 || z2 == 0) return 0;
    if (as == bs) {
    } else if (as < bs) {
        const int64_t k = sar_i64(sar_floordiv((double)(float)(bs - as), tsd));
        s1 += k < c1 ? k : c1;
        z1 = e1 - s1;
    } else if (as > bs) {
        const int64_t k = sar_i64(sar_floordiv((double)(float)(as - bs), tsd));
        s2 += k < c2 ? k : c2;
        z2 = e2 - s2;
    }
    return z1 < z2 ? z1 : z2;
}

__device__ __forceinline__ void sar_bin(int64_t b, int64_t dv, int64_t md, int64_t sz, int64_t &st, int64_t &en) {
    st = b * dv + (b < md ? b : md);
    en = (b + 1) * dv + (b + 1 < md ? b + 1 : md) + 1;
    if (en > sz) en = sz;
}

__device__ __forceinline__ bool sar_near(float2 p, float2 q, double deps) {
    const float dx = q.x - p.x, dy = q.y - p.y;
    const float d2 = dx * dx + dy * dy;
    return (double)(float)sqrt((double)d2) <= deps;
}

__device__ __forceinline__ bool sar_bin_near(double ax, double ay, double bx, double by, int64_t cnt, double deps) {
    const double cn = (double)cnt;
    ax /= cn;
    ay /= cn;
    bx /= cn;
    by /= cn;
    return sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay)) <= deps;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t w = __shfl_xor(v, o, kWave);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return v;
}

// ---- data-parallel path (non-decreasing timestamps) -----------------------

// One stretch: points [is, ie] of a trajectory, samples on grid indices
// [g0, g1) plus the last point when fin.  gl, when set, holds
// ceil(t_k / ts) - g0 for the segments k in [is, ie) (staged in LDS).
struct SarSpan {
    const double *xy, *t;
    int64_t is, ie, g0, g1;
    bool fin;
    const int32_t *gl;
    __device__ int64_t count() const { return g1 - g0 + (fin ? 1 : 0); }
    __device__ float first_time(int64_t ts) const { return g1 > g0 ? (float)(g0 * ts) : (float)t[ie]; }
    __device__ float last_time(int64_t ts) const { return fin ? (float)t[ie] : (float)((g1 - 1) * ts); }
    // sample m < count(): the segment k in [is, ie) with
    // ceil(t_k / ts) <= g < ceil(t_k+1 / ts) holds grid index g = g0 + m
    __device__ float2 sample(int64_t m, int64_t ts, double tsd) const {
        const int64_t g = g0 + m;
        if (g >= g1) return make_float2((float)xy[2 * ie], (float)xy[2 * ie + 1]);
        int64_t lo = is, hi = ie - 1;
        if (gl) {
            const int64_t gr = g - g0;
            while (lo < hi) {
                const int64_t mid = lo + (hi - lo + 1) / 2;
                if (gl[mid - is] <= gr) lo = mid;
                else hi = mid - 1;
            }
        } else {
            while (lo < hi) {
                const int64_t mid = lo + (hi - lo + 1) / 2;
                if (sar_grid(t[mid], tsd) <= g) lo = mid;
                else hi = mid - 1;
            }
        }
        return sar_interp(xy, t, lo, g * ts);
    }
};

// ---- sequential path (lane 0; any timestamps) -----------------------------

// The samples of points [k, ie] in order, as align_to_timestamps produces
// them (core/sar.py:31-59).
struct SarWalk {
    const double *xy, *t;
    int64_t k, ie, g, e;
    bool fin;
    __device__ void seek_edge(double tsd) {  // from edge k on, the next one with samples
        for (; k < ie; ++k) {
            const int64_t s = sar_grid(t[k], tsd), f = sar_grid(t[k + 1], tsd);
            if (s < f) {
                g = s;
                e = f;
                return;
            }
        }
    }
    __device__ void init(const double *xy_, const double *t_, int64_t is, int64_t ie_, double tsd) {
        xy = xy_;
        t = t_;
        k = is;
        ie = ie_;
        g = e = 0;
        fin = fmod(t[ie], tsd) == 0.0;
        seek_edge(tsd);
    }
    __device__ float2 next(int64_t ts, double tsd) {  // caller keeps within the sample count
        if (k < ie) {
            const float2 v = sar_interp(xy, t, k, g * ts);
            if (++g >= e) {
                ++k;
                seek_edge(tsd);
            }
            return v;
        }
        fin = false;
        return make_float2((float)xy[2 * ie], (float)xy[2 * ie + 1]);
    }
    __device__ void skip(int64_t n, double tsd) {
        while (n > 0 && k < ie) {
            const int64_t avail = e - g;
            if (avail > n) {
                g += n;
                return;
            }
            n -= avail;
            ++k;
            seek_edge(tsd);
        }
        if (n > 0) fin = false;
    }
};

// Sample count and first / last sample times of points [is, ie]; false when
// the count passes kSarMaxSamples.
__device__ bool sar_walk_stats(const double *t, int64_t is, int64_t ie, double tsd, int64_t ts, int64_t &cnt,
                               float &first, float &last) {
    cnt = 0;
    bool any = false;
    for (int64_t k = is; k < ie; ++k) {
        const int64_t s = sar_grid(t[k], tsd), f = sar_grid(t[k + 1], tsd);
        if (s >= f) continue;
        if (f - s > kSarMaxSamples - cnt) return false;
        if (!any) first = (float)(s * ts);
        any = true;
        last = (float)((f - 1) * ts);
        cnt += f - s;
    }
    if (fmod(t[ie], tsd) == 0.0) {
        if (!any) first = (float)t[ie];
        last = (float)t[ie];
        ++cnt;
    }
    return cnt <= kSarMaxSamples;
}

__device__ double sar_serial(const SArgs &a, const double *xa, const double *ta, int64_t na, const double *xb,
                             const double *tb, int64_t nb) {
    const double tsd = (double)a.ts;
    int64_t i = 0, j = 0, is = -1, ie = -1, js = -1, je = -1;
    while (i < na - 1 && j < nb - 1) {
        const double a0 = ta[i], a1 = ta[i + 1], b0 = tb[j], b1 = tb[j + 1];
        if (!(a1 <= b0 || b1 <= a0 || a0 == a1 || b0 == b1)) {
            if (is < 0) is = i;
            if (js < 0) js = j;
            ie = i + 1;
            je = j + 1;
        }
        if (a1 < b1) ++i;
        else if (b1 < a1) ++j;
        else {
            ++i;
            ++j;
        }
    }
    if (is < 0) return 0.0;
    int64_t c1, c2;
    float as = 0.f, ae = 0.f, bs = 0.f, be = 0.f;
    if (!sar_walk_stats(ta, is, ie, tsd, a.ts, c1, as, ae) || !sar_walk_stats(tb, js, je, tsd, a.ts, c2, bs, be))
        return sar_nan();
    if (c1 == 0 || c2 == 0) return 0.0;
    int64_t s1, s2;
    const int64_t sz = sar_trim(c1, c2, as, ae, bs, be, tsd, s1, s2);
    if (sz <= 0) return 0.0;
    SarWalk wa, wb;
    wa.init(xa, ta, is, ie, tsd);
    wb.init(xb, tb, js, je, tsd);
    wa.skip(s1, tsd);
    wb.skip(s2, tsd);
    if (a.M >= sz) {
        int64_t cnt = 0;
        for (int64_t m = 0; m < sz; ++m) {
            const float2 p = wa.next(a.ts, tsd), q = wb.next(a.ts, tsd);
            cnt += sar_near(p, q, a.deps) ? 1 : 0;
        }
        return (double)(a.R * cnt * a.ts);
    }
    const int64_t dv = sz / a.M, md = sz % a.M;
    int64_t total = 0, b = 0, st, en;
    sar_bin(0, dv, md, sz, st, en);
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    for (int64_t m = 0; m < sz; ++m) {
        const float2 p = wa.next(a.ts, tsd), q = wb.next(a.ts, tsd);
        ax += p.x;
        ay += p.y;
        bx += q.x;
        by += q.y;
        while (b < a.M && m == en - 1) {  // bin b complete; the next one starts at this sample
            if (sar_bin_near(ax, ay, bx, by, en - st, a.deps)) total += (en - st) * a.ts;
            if (++b >= a.M) break;
            sar_bin(b, dv, md, sz, st, en);
            ax = ay = bx = by = 0.0;
            if (st == m) {
                ax += p.x;
                ay += p.y;
                bx += q.x;
                by += q.y;
            }
        }
    }
    return (double)total;
}

__global__ __launch_bounds__(kWave) void sar_kernel(SArgs a) {
    __shared__ float4 chunk[kWave];  // (A.x, A.y, B.x, B.y) of 64 consecutive samples
    __shared__ int32_t gla[kSarStage], glb[kSarStage];
    const int lane = threadIdx.x;
    const double tsd = (double)a.ts;
    for (int64_t p = blockIdx.x; p < a.npairs; p += gridDim.x) {
        int64_t ia, ib;
        decode_pair(a.ps, p, ia, ib);
        if (ia < 0 || ia >= a.nset || ib < 0 || ib >= a.nset) {  // index outside the set: NaN, no access
            if (lane == 0) a.out[p] = sar_nan();
            continue;
        }
        const int64_t oa = a.offs[ia], ob = a.offs[ib];
        const int64_t na = a.offs[ia + 1] - oa, nb = a.offs[ib + 1] - ob;
        const double *ta = a.t + oa, *tb = a.t + ob;
        const double *xa = a.xy + 2 * oa, *xb = a.xy + 2 * ob;
        const double A0 = ta[0], A1 = ta[na - 1], B0 = tb[0], B1 = tb[nb - 1];

        // segments overlapping the other trajectory's span; monotonicity
        bool mono = true;
        int64_t is = INT64_MAX, ie = -1, js = INT64_MAX, je = -1;
        for (int64_t k = lane; k + 1 < na; k += kWave) {
            const double t0 = ta[k], t1 = ta[k + 1];
            mono = mono && t1 >= t0;
            if (t0 < t1 && t0 < B1 && t1 > B0 && B0 < B1) {
                is = is < k ? is : k;
                ie = k + 1;
            }
        }
        for (int64_t k = lane; k + 1 < nb; k += kWave) {
            const double t0 = tb[k], t1 = tb[k + 1];
            mono = mono && t1 >= t0;
            if (t0 < t1 && t0 < A1 && t1 > A0 && A0 < A1) {
                js = js < k ? js : k;
                je = k + 1;
            }
        }
        double res;
        if (__builtin_amdgcn_ballot_w64(!mono) != 0) {
            res = 0.0;
            if (lane == 0) res = sar_serial(a, xa, ta, na, xb, tb, nb);
        } else {
            is = wave_min_i64(is);
            ie = wave_max_i64(ie);
            js = wave_min_i64(js);
            je = wave_max_i64(je);
            res = 0.0;
            if (ie >= 0 && je >= 0) {
                SarSpan SA{xa, ta, is, ie, sar_grid(ta[is], tsd), sar_grid(ta[ie], tsd), fmod(ta[ie], tsd) == 0.0,
                           nullptr};
                SarSpan SB{xb, tb, js, je, sar_grid(tb[js], tsd), sar_grid(tb[je], tsd), fmod(tb[je], tsd) == 0.0,
                           nullptr};
                // grid indices of both stretches' segments into LDS (coalesced, one division
                // each) for the per-sample binary searches; longer stretches search in memory
                __syncthreads();  // the previous pair's readers are done
                if (ie - is <= kSarStage) {
                    for (int64_t k = lane; k < ie - is; k += kWave)
                        gla[k] = (int32_t)(sar_grid(ta[is + k], tsd) - SA.g0);
                    SA.gl = gla;
                }
                if (je - js <= kSarStage) {
                    for (int64_t k = lane; k < je - js; k += kWave)
                        glb[k] = (int32_t)(sar_grid(tb[js + k], tsd) - SB.g0);
                    SB.gl = glb;
                }
                __syncthreads();
                const int64_t c1 = SA.count(), c2 = SB.count();
                int64_t s1 = 0, s2 = 0, sz = 0;
                if (c1 > kSarMaxSamples || c2 > kSarMaxSamples) res = sar_nan();
                else if (c1 > 0 && c2 > 0)
                    sz = sar_trim(c1, c2, SA.first_time(a.ts), SA.last_time(a.ts), SB.first_time(a.ts),
                                  SB.last_time(a.ts), tsd, s1, s2);
                if (sz > 0 && a.M >= sz) {
                    int64_t cnt = 0;
                    for (int64_t c0 = 0; c0 < sz; c0 += kWave) {
                        const int64_t m = c0 + lane;
                        bool hit = false;
                        if (m < sz) hit = sar_near(SA.sample(s1 + m, a.ts, tsd), SB.sample(s2 + m, a.ts, tsd), a.deps);
                        cnt += __builtin_popcountll(__builtin_amdgcn_ballot_w64(hit));
                    }
                    res = (double)(a.R * cnt * a.ts);
                } else if (sz > 0) {
                    // lane l owns bins [l*per, (l+1)*per); sums in sample order
                    const int64_t dv = sz / a.M, md = sz % a.M, per = (a.M + kWave - 1) / kWave;
                    int64_t b = lane * per;
                    const int64_t bend = b + per < a.M ? b + per : a.M;
                    int64_t st = 0, en = 0, total = 0;
                    if (b < bend) sar_bin(b, dv, md, sz, st, en);
                    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
                    for (int64_t c0 = 0; c0 < sz; c0 += kWave) {
                        const int64_t m = c0 + lane;
                        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (m < sz) {
                            const float2 u = SA.sample(s1 + m, a.ts, tsd), w = SB.sample(s2 + m, a.ts, tsd);
                            v = make_float4(u.x, u.y, w.x, w.y);
                        }
                        __syncthreads();
                        chunk[lane] = v;
                        __syncthreads();
                        const int64_t cend = c0 + kWave < sz ? c0 + kWave : sz;
                        while (b < bend && st < cend) {
                            const int64_t from = st > c0 ? st : c0, to = en < cend ? en : cend;
                            for (int64_t k = from; k < to; ++k) {
                                const float4 q = chunk[k - c0];
                                ax += q.x;
                                ay += q.y;
                                bx += q.z;
                                by += q.w;
                            }
                            if (en > cend) break;
                            if (sar_bin_near(ax, ay, bx, by, en - st, a.deps)) total += (en - st) * a.ts;
                            ax = ay = bx = by = 0.0;
                            if (++b < bend) sar_bin(b, dv, md, sz, st, en);
                        }
                    }
                    res = (double)wave_sum_i64(total);
                }
            }
        }
        if (lane == 0) a.out[p] = res;
    }
}

}  // namespace tsm
