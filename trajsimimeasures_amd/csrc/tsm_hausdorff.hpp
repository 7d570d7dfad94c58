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
// time (128, two per lane, while more than 64 remain); the other side's
// segments are wave-uniform, so their endpoints and
// per-segment constants (dx, dy, 1/|s|^2, computed once per call by
// seg_prep_kernel) come through scalar loads and every lane keeps a running
// minimum in a VGPR.  A wave max-reduction and one square root finish the
// pair.
#pragma once

#include "tsm_device.hpp"

namespace tsm {

struct HArgs {
    const float *xy32;
    const double *xy64;
    const float4 *seg;  // FP32 path: per point i, segment (p_i, p_i+1): (x_i+1, y_i+1, dx, dy)
    const float *rseg;  // FP32 path: 1/|s|^2, or 0 for a degenerate segment
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

// FP32 form with precomputed segment constants.  G[k] = (x, y) of the
// segment's end point s2 and (dx, dy) = s2 - s1; R[k] = 1/|s2 - s1|^2, or 0
// for a degenerate segment.  One clamped projection gives the reference's
// value in every case: with c = 0 if u < 1e-5, else min(u, 1), the residual
// |(p - s1) - c (s2 - s1)|^2 is
//   - |p - s1|^2 exactly when u < 1e-5 (c = 0; the reference's
//     min(dps1, dps2) is dps1 there, since u < 1/2 means s1 is nearer), and
//     for a degenerate segment (R = 0 gives u = 0), as the reference;
//   - |p - s2|^2 up to rounding when u > 1 (c = 1; u > 1 means s2 is
//     nearer, so min(dps1, dps2) = dps2);
//   - the distance to the foot of the perpendicular otherwise.
// Both arrays are read with wave-uniform addresses (scalar loads).
struct SegState {
    float ex, ey, best;  // p - s1 of the current segment, running minimum
};

__device__ __forceinline__ void p2s_step(float px, float py, float4 g, float r, SegState &st) {
    const float u = fmaf(st.ey, g.w, st.ex * g.z) * r;
    const float c = u < 0.00001f ? 0.0f : fminf(u, 1.0f);
    const float gx = fmaf(-c, g.z, st.ex), gy = fmaf(-c, g.w, st.ey);
    st.best = fminf(st.best, fmaf(gy, gy, gx * gx));
    st.ex = px - g.x;  // p - s2 is the next segment's p - s1
    st.ey = py - g.y;
}

__device__ __forceinline__ float p2t_sq_f32(float px, float py, const float *__restrict__ Y,
                                            const float4 *__restrict__ G, const float *__restrict__ R, int ny) {
    SegState st;
    st.ex = px - Y[0];
    st.ey = py - Y[1];
    st.best = __builtin_inff();
    const int nseg = ny - 1;
    // G and R are indexed by the wave-uniform k: read them as scalars
    const int *Gi = reinterpret_cast<const int *>(G);
    auto uni = [](int v) { return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(v)); };
    auto seg_at = [&](int q) { return make_float4(uni(Gi[4 * q]), uni(Gi[4 * q + 1]), uni(Gi[4 * q + 2]), uni(Gi[4 * q + 3])); };
    int k = 0;
    for (; k + 8 <= nseg; k += 8) {  // eight segments per scalar-load wait
        float4 g[8];
        float r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            g[j] = seg_at(k + j);
            r[j] = R[k + j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) p2s_step(px, py, g[j], r[j], st);
    }
    for (; k + 4 <= nseg; k += 4) {
        const float4 g0 = seg_at(k), g1 = seg_at(k + 1), g2 = seg_at(k + 2), g3 = seg_at(k + 3);
        p2s_step(px, py, g0, R[k], st);
        p2s_step(px, py, g1, R[k + 1], st);
        p2s_step(px, py, g2, R[k + 2], st);
        p2s_step(px, py, g3, R[k + 3], st);
    }
    for (; k < nseg; ++k) p2s_step(px, py, seg_at(k), R[k], st);
    return st.best;
}

// The same for two points per lane: each segment's scalar loads feed twice
// the arithmetic (used when more than a wavefront of points remains).
__device__ __forceinline__ void p2t_sq_f32_x2(float px0, float py0, float px1, float py1, const float *__restrict__ Y,
                                              const float4 *__restrict__ G, const float *__restrict__ R, int ny,
                                              float &d0, float &d1) {
    SegState s0, s1;
    s0.ex = px0 - Y[0];
    s0.ey = py0 - Y[1];
    s1.ex = px1 - Y[0];
    s1.ey = py1 - Y[1];
    s0.best = s1.best = __builtin_inff();
    const int nseg = ny - 1;
    const int *Gi = reinterpret_cast<const int *>(G);
    auto uni = [](int v) { return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(v)); };
    auto seg_at = [&](int q) { return make_float4(uni(Gi[4 * q]), uni(Gi[4 * q + 1]), uni(Gi[4 * q + 2]), uni(Gi[4 * q + 3])); };
    int k = 0;
    for (; k + 4 <= nseg; k += 4) {  // four segments per scalar-load wait
        const float4 g0 = seg_at(k), g1 = seg_at(k + 1), g2 = seg_at(k + 2), g3 = seg_at(k + 3);
        const float r0 = R[k], r1 = R[k + 1], r2 = R[k + 2], r3 = R[k + 3];
        p2s_step(px0, py0, g0, r0, s0);
        p2s_step(px1, py1, g0, r0, s1);
        p2s_step(px0, py0, g1, r1, s0);
        p2s_step(px1, py1, g1, r1, s1);
        p2s_step(px0, py0, g2, r2, s0);
        p2s_step(px1, py1, g2, r2, s1);
        p2s_step(px0, py0, g3, r3, s0);
        p2s_step(px1, py1, g3, r3, s1);
    }
    for (; k < nseg; ++k) {
        const float4 g0 = seg_at(k);
        const float r0 = R[k];
        p2s_step(px0, py0, g0, r0, s0);
        p2s_step(px1, py1, g0, r0, s1);
    }
    d0 = s0.best;
    d1 = s1.best;
}

// Per point i of every trajectory: segment (p_i, p_i+1) as G[i] = (x_i+1,
// y_i+1, dx, dy) and R[i] = 1/|s|^2 rounded up (0 if degenerate); the last
// point of a trajectory gets zeros (never read).
__global__ void seg_prep_kernel(const float *xy, int64_t npts, float4 *G, float *R) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npts; i += (int64_t)gridDim.x * blockDim.x) {
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        float r = 0.f;
        if (i + 1 < npts) {
            const float x1 = xy[2 * i], y1 = xy[2 * i + 1], x2 = xy[2 * i + 2], y2 = xy[2 * i + 3];
            const float xd = x2 - x1, yd = y2 - y1;
            g = make_float4(x2, y2, xd, yd);
            if (!(x1 == x2 && y1 == y2)) {
                // 1/|s|^2 rounded up, so that r |s|^2 >= 1 exactly: a point
                // equal to s2 has u1 = |s|^2 (the same expression on the same
                // values), hence u >= 1, c = 1 and a residual of exactly 0,
                // as the reference's FP64 u = 1 gives
                const float segl = fmaf(yd, yd, xd * xd);
                r = 1.0f / segl;
                if (fmaf(r, segl, -1.0f) < 0.0f)  // r > 0 finite: the next float up
                    r = __builtin_bit_cast(float, __builtin_bit_cast(unsigned int, r) + 1u);
            }
        }
        G[i] = g;
        R[i] = r;
    }
}

template <typename T>
__global__ __launch_bounds__(kWave) void hausdorff_kernel(HArgs a) {
    const int lane = threadIdx.x;
    const T *xy = HPt<T>::base(a);
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
            for (int base = 0; base < nx;) {
                if constexpr (sizeof(T) == 4) {
                    const float4 *G = a.seg + (dir == 0 ? ob : oa);
                    const float *R = a.rseg + (dir == 0 ? ob : oa);
                    if (nx - base > kWave) {  // wave-uniform: points base + lane and base + 64 + lane
                        const int i0 = base + lane;
                        int i1 = base + kWave + lane;
                        const bool ok1 = i1 < nx;
                        i1 = ok1 ? i1 : nx - 1;
                        float d0, d1;
                        p2t_sq_f32_x2(X[2 * i0], X[2 * i0 + 1], X[2 * i1], X[2 * i1 + 1], Y, G, R, ny, d0, d1);
                        m = d0 > m ? d0 : m;
                        if (ok1) m = d1 > m ? d1 : m;
                        base += 2 * kWave;
                        continue;
                    }
                }
                int i = base + lane;
                const bool ok = i < nx;
                i = ok ? i : nx - 1;
                T d;
                if constexpr (sizeof(T) == 4)
                    d = p2t_sq_f32(X[2 * i], X[2 * i + 1], Y, a.seg + (dir == 0 ? ob : oa),
                                   a.rseg + (dir == 0 ? ob : oa), ny);
                else
                    d = p2t_sq<T>(X[2 * i], X[2 * i + 1], Y, ny);
                if (ok) m = d > m ? d : m;
                base += kWave;
            }
        }
        double r;
        if constexpr (sizeof(T) == 4) r = (double)wave_maxf(m);
        else r = wave_maxd(m);
        if (lane == 0) a.out[p] = sqrt(r);
    }
}

}  // namespace tsm
