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
// Mapping: one wavefront per pair.  Lanes take 64 points of one side at a
// time; the other side's segments are wave-uniform, so their endpoints come
// through scalar loads and every lane keeps a running minimum in a VGPR.  A
// wave max-reduction and one square root finish the pair.
#pragma once

#include "tsm_device.hpp"

namespace tsm {

struct HArgs {
    const float *xy32;
    const double *xy64;
    const int64_t *offs;
    PairSpace ps;
    int64_t npairs;
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

template <typename T>
__global__ __launch_bounds__(kWave) void hausdorff_kernel(HArgs a) {
    const int lane = threadIdx.x;
    const T *xy = HPt<T>::base(a);
    for (int64_t p = blockIdx.x; p < a.npairs; p += gridDim.x) {
        int64_t ia, ib;
        decode_pair(a.ps, p, ia, ib);
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
            for (int base = 0; base < nx; base += kWave) {
                int i = base + lane;
                const bool ok = i < nx;
                i = ok ? i : nx - 1;
                const T d = p2t_sq<T>(X[2 * i], X[2 * i + 1], Y, ny);
                if (ok) m = d > m ? d : m;
            }
        }
        double r;
        if constexpr (sizeof(T) == 4) r = (double)wave_maxf(m);
        else r = wave_maxd(m);
        if (lane == 0) a.out[p] = sqrt(r);
    }
}

}  // namespace tsm
