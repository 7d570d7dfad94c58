// tsm_hausdorff.hpp -- symmetric point-to-segment Hausdorff distance
// (core/hausdorff.py:12-123).
//
// h(A, B) = max over points p of A of min over segments (b_k, b_k+1) of B of
// the squared point-to-segment distance, with the reference's projection
// rule: a degenerate segment uses |p - s1|^2 (hausdorff.py:29-30); otherwise
// u = ((p - s1) . (s2 - s1)) / |s2 - s1|^2 and u < 1e-5 or u > 1 uses the
// nearer endpoint (hausdorff.py:37-39), else the foot of the perpendicular.
// A polyline with one point has no segment and contributes 0
// (hausdorff.py:50-51).  The result is sqrt(max(h(A,B), h(B,A))), with padding
// excluded from the max.
//
// Mapping: one wavefront per pair.  For each direction the wave computes
// the other side's segment constants (start point, direction, direction / |s|^2)
// into LDS, then lanes hold up to 8 points each and sweep a share of the
// segments, keeping running minima in VGPRs (hd_layout); the minima meet
// across lane groups, a wave max-reduction and one square root finish the
// pair.  Nothing is precomputed per call over the whole set.
#pragma once

#include "tsm_device.hpp"

namespace tsm {

struct HArgs {
    const float *xy32;
    const double *xy64;
    const int64_t *offs;
    PairSpace ps;
    int64_t npairs;
    int64_t nset;  // trajectories in the set
    double *out;
};

template <typename T> struct HPt;
template <> struct HPt<float> {
    __device__ static const float *base(const HArgs &a) { return a.xy32; }
};
template <> struct HPt<double> {
    __device__ static const double *base(const HArgs &a) { return a.xy64; }
};

// a*b + c: one FMA in FP32; in FP64 a rounded product then a rounded sum,
// the oracle's operation order.
__device__ __forceinline__ float madd(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double madd(double a, double b, double c) { return a * b + c; }

// Squared distance from (px, py) to polyline Y[0..ny) (ny >= 2); the running
// minimum over segments of the reference's p2s_dist (hausdorff.py:12-69).
template <typename T>
__device__ __forceinline__ T p2t_sq(T px, T py, const T *__restrict__ Y, int ny) {
    T s1x = Y[0], s1y = Y[1];
    T ex = px - s1x, ey = py - s1y;
    T dps1 = madd(ey, ey, ex * ex);
    T best = (T)__builtin_inf();
    for (int k = 0; k + 1 < ny; ++k) {
        const T s2x = Y[2 * k + 2], s2y = Y[2 * k + 3];
        const T fx = px - s2x, fy = py - s2y;
        const T dps2 = madd(fy, fy, fx * fx);
        T dist;
        if (s1x == s2x && s1y == s2y) {
            dist = dps1;
        } else {
            const T xd = s2x - s1x, yd = s2y - s1y;
            const T segl = madd(yd, yd, xd * xd);
            const T u1 = madd(ey, yd, ex * xd);
            T u;
            if constexpr (sizeof(T) == 4) u = u1 * __builtin_amdgcn_rcpf(segl);
            else u = u1 / segl;
            if (u < (T)0.00001 || u > (T)1) {
                dist = dps1 < dps2 ? dps1 : dps2;
            } else {
                const T ix = madd(u, xd, s1x), iy = madd(u, yd, s1y);
                const T gx = px - ix, gy = py - iy;
                dist = madd(gy, gy, gx * gx);
            }
        }
        best = dist < best ? dist : best;
        s1x = s2x;
        s1y = s2y;
        ex = fx;
        ey = fy;
        dps1 = dps2;
    }
    return best;
}

// FP32 form.  Per segment k of Y the wave stages in LDS its start point and
// direction G[k] = (x1, y1, dx, dy) and Q[k] = (dx, dy) r', with r' the
// reciprocal of |s|^2 rounded up and widened by 2^-20 (0 for a degenerate
// segment).  One clamped projection gives the reference's value in every
// case: with u = (p - s1) . Q and c = 0 if u < 1e-5, else min(u, 1), the
// residual |(p - s1) - c (s2 - s1)|^2 is
//   - |p - s1|^2 exactly when u < 1e-5 (c = 0; the reference's
//     min(dps1, dps2) is dps1 there, since u < 1/2 means s1 is nearer), and
//     for a degenerate segment (Q = 0 gives u = 0), as the reference;
//   - |p - s2|^2 up to rounding when u > 1 (c = 1; u > 1 means s2 is
//     nearer, so min(dps1, dps2) = dps2);
//   - the distance to the foot of the perpendicular otherwise.
// A point equal to s2 has p - s1 = (dx, dy) bit for bit, and then u >= 1:
// with rounding unit e = 2^-24, r |s|^2 >= 1 for the computed |s|^2, which
// is at most (dx^2 + dy^2)(1 + e)^2, and the two products and the FMA of u
// lose at most (1 - e)^3, so u >= (1 + 2^-20)(1 - e)^6 / (1 + e)^2 > 1:
// c = 1 and the residual is exactly 0, as the reference's FP64 u = 1
// gives.  Elsewhere the widening moves u by at most 2^-20 relative, which
// moves the foot of the perpendicular along the segment and the squared
// distance by a second-order amount.
__device__ __forceinline__ void p2s_eval(float px, float py, float4 g, float2 q, float &best) {
    const float ex = px - g.x, ey = py - g.y;
    const float u = fmaf(ey, q.y, ex * q.x);
    const float c = u < 0.00001f ? 0.0f : fminf(u, 1.0f);
    const float gx = fmaf(-c, g.z, ex), gy = fmaf(-c, g.w, ey);
    best = fminf(best, fmaf(gy, gy, gx * gx));
}

__device__ __forceinline__ void seg_consts(float x1, float y1, float x2, float y2, float4 &g, float2 &q) {
    const float xd = x2 - x1, yd = y2 - y1;
    g = make_float4(x1, y1, xd, yd);
    q = make_float2(0.0f, 0.0f);
    if (!(x1 == x2 && y1 == y2)) {
        const float segl = fmaf(yd, yd, xd * xd);
        float r = 1.0f / segl;
        if (fmaf(r, segl, -1.0f) < 0.0f)  // r > 0 finite: the next float up
            r = __builtin_bit_cast(float, __builtin_bit_cast(unsigned int, r) + 1u);
        r *= 1.00000095367431640625f;  // 1 + 2^-20
        q = make_float2(xd * r, yd * r);
    }
}

// Segments staged per LDS chunk: 24 B each, 19,968 B per wave, so 8 waves
// share a CU.
constexpr int kHChunk = 832;

// Lane layout of one direction (points of X against the segments of Y):
// Ps groups of Lg = 64 / Ps lanes; lane t of group g holds points
// t, t + Lg, ..., t + (M - 1) Lg of each pass of M Lg points, and group g
// takes segments g, g + Ps, ...; the minima meet across the groups at the
// end of a pass.  (M, Ps) is the pair with the fewest lane-steps,
// M * ceil(nseg / Ps) with M Lg >= nx, M <= 8: on lengths 20..200 the
// lanes do useful work 91 % of the time, against 78 % for passes of 64 or
// 128 points over every segment.
__device__ __forceinline__ void hd_layout(int nx, int nseg, int &M, int &Ps) {
    M = 8;
    Ps = 1;
    int best = 1 << 30;
    for (int ps = 1; ps <= 16; ps <<= 1) {
        const int lg = kWave / ps;
        const int m = (nx + lg - 1) / lg;
        if (m > 8) continue;
        const int cost = m * ((nseg + ps - 1) / ps);
        if (cost < best) {
            best = cost;
            M = m;
            Ps = ps;
        }
    }
}

// Largest over X's points of the squared distance to polyline Y (ny >= 2):
// this lane's share (the wave max of the results is the direction's value).
template <int M>
__device__ __forceinline__ float hd_dir_f32(const float *__restrict__ X, int nx, const float *__restrict__ Y, int ny,
                                         int Ps, float4 *sg, float2 *sq) {
    const int lane = threadIdx.x;
    const int Lg = kWave / Ps;
    const int grp = lane / Lg, t = lane % Lg;
    const int nseg = ny - 1;
    float m = 0.0f;
    for (int base = 0; base < nx; base += M * Lg) {
        float px[M], py[M], best[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            int i = base + t + q * Lg;
            i = i < nx ? i : nx - 1;
            px[q] = X[2 * i];
            py[q] = X[2 * i + 1];
            best[q] = __builtin_inff();
        }
        for (int c0 = 0; c0 < nseg; c0 += kHChunk) {
            const int nc = nseg - c0 < kHChunk ? nseg - c0 : kHChunk;
            __syncthreads();  // the previous chunk's reads are done
            for (int k = lane; k < nc; k += kWave) {
                const float *s = Y + 2 * (c0 + k);
                seg_consts(s[0], s[1], s[2], s[3], sg[k], sq[k]);
            }
            __syncthreads();
#pragma unroll 4
            for (int k = grp; k < nc; k += Ps) {
                const float4 g = sg[k];
                const float2 d = sq[k];
#pragma unroll
                for (int q = 0; q < M; ++q) p2s_eval(px[q], py[q], g, d, best[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < M; ++q) {
            for (int o = Lg; o < kWave; o <<= 1) best[q] = fminf(best[q], __shfl_xor(best[q], o, kWave));
            if (base + t + q * Lg < nx) m = fmaxf(m, best[q]);
        }
    }
    return m;
}

__device__ __forceinline__ float hd_dir_dispatch(const float *X, int nx, const float *Y, int ny, float4 *sg,
                                                 float2 *sq) {
    int M, Ps;
    hd_layout(nx, ny - 1, M, Ps);
    switch (M) {
    case 1: return hd_dir_f32<1>(X, nx, Y, ny, Ps, sg, sq);
    case 2: return hd_dir_f32<2>(X, nx, Y, ny, Ps, sg, sq);
    case 3: return hd_dir_f32<3>(X, nx, Y, ny, Ps, sg, sq);
    case 4: return hd_dir_f32<4>(X, nx, Y, ny, Ps, sg, sq);
    case 5: return hd_dir_f32<5>(X, nx, Y, ny, Ps, sg, sq);
    case 6: return hd_dir_f32<6>(X, nx, Y, ny, Ps, sg, sq);
    case 7: return hd_dir_f32<7>(X, nx, Y, ny, Ps, sg, sq);
    default: return hd_dir_f32<8>(X, nx, Y, ny, Ps, sg, sq);
    }
}

template <typename T>
__global__ __launch_bounds__(kWave) void hausdorff_kernel(HArgs a) {
    const int lane = threadIdx.x;
    const T *xy = HPt<T>::base(a);
    __shared__ float4 sg[sizeof(T) == 4 ? kHChunk : 1];
    __shared__ float2 sq[sizeof(T) == 4 ? kHChunk : 1];
    for (int64_t p = blockIdx.x; p < a.npairs; p += gridDim.x) {
        int64_t ia, ib;
        decode_pair(a.ps, p, ia, ib);
        if (ia < 0 || ia >= a.nset || ib < 0 || ib >= a.nset) {  // index outside the set: NaN, no access
            if (lane == 0) a.out[p] = __builtin_bit_cast(double, 0x7ff8000000000000ull);
            continue;
        }
        const int64_t oa = a.offs[ia], ob = a.offs[ib];
        const int na = (int)(a.offs[ia + 1] - oa), nb = (int)(a.offs[ib + 1] - ob);
        T m = 0;
#pragma unroll 1
        for (int dir = 0; dir < 2; ++dir) {
            const T *X = xy + 2 * (dir == 0 ? oa : ob);
            const T *Y = xy + 2 * (dir == 0 ? ob : oa);
            const int nx = dir == 0 ? na : nb;
            const int ny = dir == 0 ? nb : na;
            if (ny < 2) continue;
            if constexpr (sizeof(T) == 4) {
                const float d = hd_dir_dispatch(X, nx, Y, ny, sg, sq);
                m = d > m ? d : m;
            } else {
                for (int base = 0; base < nx; base += kWave) {
                    int i = base + lane;
                    const bool ok = i < nx;
                    i = ok ? i : nx - 1;
                    const T d = p2t_sq<T>(X[2 * i], X[2 * i + 1], Y, ny);
                    if (ok) m = d > m ? d : m;
                }
            }
        }
        double r;
        if constexpr (sizeof(T) == 4) r = (double)wave_maxf(m);
        else r = wave_maxd(m);
        if (lane == 0) a.out[p] = sqrt(r);
    }
}

}  // namespace tsm
