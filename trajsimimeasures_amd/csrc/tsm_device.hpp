// tsm_device.hpp -- device-side building blocks shared by the DP and
// Hausdorff kernels of libtsm (gfx950 / CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsm {

constexpr int kWave = 64;
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;

enum Measure : int { kDTW = 0, kEDR = 1, kLCSS = 2, kERP = 3, kFrechet = 4, kHausdorff = 5, kSTEDR = 6, kSTLCSS = 7 };

// Pair addressing modes: explicit index arrays, a Q x D grid, or the upper
// triangle of an all-pairs matrix.
enum PairMode : int { kExplicit = 0, kGrid = 1, kTriu = 2 };

struct PairSpace {
    int mode;
    const int32_t *pa;
    const int32_t *pb;
    int64_t q0, nq, d0, nd;  // grid
    int64_t n;               // triangle side
};

// Start index of row i of the row-major strict upper triangle of an n x n matrix.
__device__ __forceinline__ int64_t triu_row_start(int64_t i, int64_t n) { return i * (2 * n - i - 1) / 2; }

// Pair index -> (t1, t2) trajectory indices.
__device__ __forceinline__ void decode_pair(const PairSpace &ps, int64_t p, int64_t &ia, int64_t &ib) {
    if (ps.mode == kExplicit) {
        ia = ps.pa[p];
        ib = ps.pb[p];
    } else if (ps.mode == kGrid) {
        int64_t q = p / ps.nd;
        ia = ps.q0 + q;
        ib = ps.d0 + (p - q * ps.nd);
    } else {
        const int64_t n = ps.n;
        double b = 2.0 * (double)n - 1.0;
        int64_t i = (int64_t)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
        if (i < 0) i = 0;
        if (i > n - 2) i = n - 2;
        while (i > 0 && triu_row_start(i, n) > p) --i;
        while (i < n - 2 && triu_row_start(i + 1, n) <= p) ++i;
        ia = i;
        ib = p - triu_row_start(i, n) + i + 1;
    }
}

// Lane i <- lane i-1 within each row of 16 lanes (DPP row_shr:1).  Lane 0 of
// each row receives 0; callers overwrite it.
__device__ __forceinline__ float shr1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF, false));
}
__device__ __forceinline__ double shr1(double v) {
    int64_t b = __builtin_bit_cast(int64_t, v);
    int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), 0x111, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x111, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

// Lane i <- lane i-1 across the whole wavefront (DPP wave_shr:1).  Lane 0
// receives 0; callers overwrite it.
__device__ __forceinline__ float shr1_wave(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double shr1_wave(double v) {
    int64_t b = __builtin_bit_cast(int64_t, v);
    int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), 0x138, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_maxd(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(fminf(a, b), c); }
__device__ __forceinline__ double min3d(double a, double b, double c) { return fmin(fmin(a, b), c); }

// Euclidean distance with one hardware square root (v_sqrt_f32, <= 1 ulp).
__device__ __forceinline__ float dist_f32(float ax, float ay, float bx, float by) {
    float dx = ax - bx, dy = ay - by;
    return __builtin_amdgcn_sqrtf(fmaf(dy, dy, dx * dx));
}
__device__ __forceinline__ float dist2_f32(float ax, float ay, float bx, float by) {
    float dx = ax - bx, dy = ay - by;
    return fmaf(dy, dy, dx * dx);
}
// FP64 squared distance with the oracle's operation order and no FMA:
// (dx*dx) + (dy*dy), each rounded (the file is built with -ffp-contract=off).
__device__ __forceinline__ double dist2_f64(double ax, double ay, double bx, double by) {
    double dx = ax - bx, dy = ay - by;
    return dx * dx + dy * dy;
}

}  // namespace tsm
