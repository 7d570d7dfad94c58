// tsm_host.hpp -- host-side state shared by the translation units of
// libtsm.so: the per-process device context, grow-only workspace buffers,
// error reporting and launch helpers.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tsm.h"
#include "tsm_device.hpp"

namespace tsm {

int fail(int code, const std::string &msg);

#define TSM_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(TSM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    inline int ensure(size_t bytes) {
        if (bytes <= cap) return TSM_OK;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return fail(TSM_E_HIP, std::string("hipFree: ") + hipGetErrorString(e));
        }
        size_t want = bytes + bytes / 4 + 256;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(TSM_E_NOMEM, "hipMalloc(" + std::to_string(want) + ") failed: " + hipGetErrorString(e));
        }
        cap = want;
        e = hipMemset(p, 0, want);  // fresh buffers start zeroed (the plan histogram relies on it)
        if (e != hipSuccess) return fail(TSM_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        return TSM_OK;
    }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Fixed device memory handed out by a bump cursor in 256-byte aligned
// pieces, in place of grow-only allocations (the no-allocation entry point).
// Its contents are whatever the previous call left there.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};

// Device scratch of one call sequence: the plan and the strip boundary
// columns.  The context owns the shared one;
// each captured graph owns its own, so replays never see reallocations.  A
// workspace over an arena allocates nothing: its buffers are pieces of the
// arena, valid for one call, and must not be released.
struct Workspace {
    DevBuf hist, slots, meta, counter, bnd;  // hist: the planner's histogram, then its fill counters
    Arena *arena = nullptr;
    // b holding at least `bytes`: grown (fresh memory is zeroed), or the next
    // piece of the arena (contents left by the previous call)
    int get(DevBuf &b, size_t bytes) {
        if (!arena) return b.ensure(bytes);
        const size_t off = (arena->used + 255) & ~(size_t)255;
        if (off > arena->cap || bytes > arena->cap - off)
            return fail(TSM_E_ARG, "the call needs more device scratch than the " + std::to_string(arena->cap >> 20) +
                                       " MiB of the no-allocation entry point (fewer pairs, points or shorter "
                                       "trajectories per call)");
        b.p = arena->base + off;
        b.cap = bytes;
        arena->used = off + bytes;
        return TSM_OK;
    }
    // bytes of the arena still free (0 without one)
    size_t arena_left() const {
        if (!arena) return 0;
        const size_t off = (arena->used + 255) & ~(size_t)255;
        return off < arena->cap ? arena->cap - off : 0;
    }
    void release() {
        if (arena) return;
        for (DevBuf *b : {&hist, &slots, &meta, &counter, &bnd})
            b->release();
    }
};

// Zero n 32-bit words (a fill kernel: no memset node in a captured graph).
int clear_words(uint32_t *p, int64_t n, hipStream_t st);

struct Context {
    int device = -1;
    int num_cu = 0;
    Workspace ws;
    // staging for tsm_pairs_host / tsm_pairs_padded_host
    DevBuf h_xy64, h_xy32, h_offs, h_pa, h_pb, h_out, h_cnt, h_t64, h_pad, h_box;
    // autotuned lanes per column count (index 0: FP32 coordinates, 1: FP64)
    uint8_t tuned_lanes[2][258] = {};
    bool tuned[2] = {false, false};
    int waves_per_cu = 8;  // resident waves per CU of the persistent DP kernel (autotuned)
    int wide_lanes = 0;    // lanes for pairs wider than the bucket table: 0 = by launch size, 16 or 64
    int64_t direct_pairs = -1;  // direct DP launches up to this many pairs: -1 = kDirectRounds * num_cu *
                                // waves_per_cu, 0 = never
    // ordering of the shared workspace across streams: the last call's
    // stream and an event recorded after its kernels
    hipEvent_t ws_event = nullptr;
    hipStream_t ws_stream = nullptr;
    bool ws_used = false;
    bool capturing = false;  // inside a graph capture or a no-allocation call: no profiling events
    // profiling: every prof_every-th timed launch gets an event pair
    bool prof = false;
    int prof_every = 1;
    int64_t prof_seen = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
};

extern std::mutex g_mu;
extern Context g_ctx;

inline int ensure_ctx() {
    int dev = 0;
    TSM_HIP(hipGetDevice(&dev));
    if (g_ctx.device == dev) return TSM_OK;
    if (g_ctx.device >= 0) return fail(TSM_E_UNSUP, "libtsm serves one device per process");
    hipDeviceProp_t prop;
    TSM_HIP(hipGetDeviceProperties(&prop, dev));
    g_ctx.device = dev;
    g_ctx.num_cu = prop.multiProcessorCount;
    return TSM_OK;
}

inline double eps_threshold(double eps, int mode) {
    if (mode == TSM_EPS_SQUARED) return eps * eps;
    if (!(eps > 0.0)) return 0.0;
    if (std::isinf(eps)) return INFINITY;
    double t = eps * eps;
    while (std::sqrt(t) >= eps) t = std::nextafter(t, 0.0);
    while (std::sqrt(t) < eps) t = std::nextafter(t, INFINITY);
    return t;
}

// Finite tests.  The library is built with -fno-honor-nans, under which
// the compiler folds isnan / isfinite as if NaN never occurred, and also
// rewrites an exponent-bits test of a double's bit pattern into a floating
// class test whose NaN half it then drops (what remains is |v| == inf).  So
// the bits must reach the test as an integer the compiler cannot trace back
// to a double: loaded from memory as uint64_t (finite_bits), or, for a
// double held in a register, passed through a volatile (finite64; host
// code, a few scalars per call).
__host__ __device__ inline bool finite_bits(uint64_t b) {
    return (uint32_t)(b >> 52 & 0x7ffu) != 0x7ffu;
}
inline bool finite64(double v) {
    volatile uint64_t b = __builtin_bit_cast(uint64_t, v);
    return finite_bits(b);
}

// Guard band of the FP32 match test of EDR/LCSS.  With u = 2^-24, U a bound
// on |coordinate| and r = sqrt(S), the FP32 squared distance d2' computed
// from FP32-rounded coordinates (two subtractions, a product and an FMA)
// differs from the FP64 oracle's d2 by at most
//   6u S + 8u U r + 8u^2 (U + r)^2 + 4e-16 S
// (coordinate rounding u|a|, subtraction rounding u|dx'|, squaring, two
// roundings of the sum).  For S near thr, r >= sqrt(thr / 2), so the error
// relative to thr is at most
//   rel = 7u + 12u U / sqrt(thr) + 20u^2 (U + sqrt(thr))^2 / thr,
// and for S far from thr the error cannot move d2' across the band.  With
// the band [thr (1 - 3 rel), thr (1 + 3 rel)) rounded outwards to FP32, a
// value below it is a match and a value at or above it is not, exactly as
// in FP64.  Returns false when no useful band exists (unknown or huge
// coordinates relative to eps): then the test runs in FP64.  A band has
// 0 <= lo <= hi (the kernel orders FP32 bit patterns, which holds only for
// non-negative values).
// coord_max bounds the FP32 copy's coordinates, which may be taken relative
// to an origin: fl32(fl64(x - o)).  The FP64 subtraction adds at most
// 2^-53 |x - o| to the u |x - o| of the FP32 rounding; U is widened by
// 2^-20 to cover it (u 2^-20 U = 2^-44 U > 2^-52 U).
inline bool count_band(double thr, double coord_max, float &lo, float &hi) {
    if (!finite64(coord_max) || !(coord_max > 0.0) || coord_max > 1e15) return false;
    if (!(thr > 0.0)) {  // nothing matches
        lo = hi = 0.0f;
        return true;
    }
    if (std::isinf(thr)) {  // everything matches (finite FP32 squared distances)
        lo = hi = INFINITY;
        return true;
    }
    const double u = 0x1p-24, e = std::sqrt(thr), U = coord_max * (1.0 + 0x1p-20);
    const double rel = 7.0 * u + 12.0 * u * U / e + 20.0 * u * u * (U + e) * (U + e) / thr;
    const double B = 3.0 * rel;
    if (!(B <= 1e-2)) return false;
    const double dlo = thr * (1.0 - B), dhi = thr * (1.0 + B);
    lo = (float)dlo;
    if ((double)lo > dlo) lo = std::nextafter(lo, 0.0f);
    hi = (float)dhi;
    if ((double)hi < dhi) hi = std::nextafter(hi, INFINITY);
    return true;
}

// Origin of a set's FP32 copy: per axis, the middle of the bounding box
// rounded to a multiple of the FP32 spacing of the axis' largest |value|.
// Every FP32 value on the axis is a multiple of its own spacing, which
// divides that one, so x - o is exact in FP32 whenever |x - o| stays below
// 2^24 spacings of x (FP32 inputs of one binade per axis, as Mercator
// coordinates are).  cmax bounds |x - ox| and |y - oy|.  From the box
// lo[2], hi[2] of a non-empty set of finite points:
inline void fp32_origin_of_box(const double lo[2], const double hi[2], double &ox, double &oy, double &cmax) {
    double o[2] = {0.0, 0.0};
    cmax = 0.0;
    for (int k = 0; k < 2; ++k) {
        const double big = std::fmax(std::fabs(lo[k]), std::fabs(hi[k]));
        if (big > 0.0) {
            const double q = std::ldexp(1.0, std::ilogb(big) - 23);  // FP32 spacing at |big|
            o[k] = std::nearbyint((0.5 * lo[k] + 0.5 * hi[k]) / q) * q;
        }
        cmax = std::fmax(cmax, std::fmax(std::fabs(lo[k] - o[k]), std::fabs(hi[k] - o[k])));
    }
    ox = o[0];
    oy = o[1];
}

// The same from the points xy [npts][2]; returns nonzero when a coordinate
// is not finite (or npts < 1).
inline int fp32_origin(const double *xy, int64_t npts, double &ox, double &oy, double &cmax) {
    if (npts < 1) return 1;
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    const uint64_t *bits = reinterpret_cast<const uint64_t *>(xy);
    for (int64_t i = 0; i < npts; ++i)
        for (int k = 0; k < 2; ++k) {
            const uint64_t b = bits[2 * i + k];
            if (!finite_bits(b)) return 1;
            const double v = __builtin_bit_cast(double, b);
            lo[k] = std::fmin(lo[k], v);
            hi[k] = std::fmax(hi[k], v);
        }
    fp32_origin_of_box(lo, hi, ox, oy, cmax);
    return 0;
}

inline int grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

// Grid of a persistent kernel: waves_per_cu waves per CU (autotuned, 8 by
// default; the TSM_WAVES_PER_CU environment variable overrides it), or fewer
// when there is less work.  More waves than the registers allow wait for a
// slot and exit at once when they find the queue drained.  Measured: beyond
// 2 resident waves per SIMD the extra waves add memory-wait time on the row
// loads faster than they hide it.
inline int persistent_grid(int64_t max_useful) {
    static const int env = [] {
        const char *e = std::getenv("TSM_WAVES_PER_CU");
        return e ? std::atoi(e) : 0;
    }();
    const int waves_per_cu = env > 0 ? env : g_ctx.waves_per_cu;
    int64_t g = (int64_t)g_ctx.num_cu * waves_per_cu;
    if (g > max_useful) g = max_useful;
    return (int)(g < 1 ? 1 : g);
}

// Kernel timing (tsm_profile_enable): a start / stop event pair for
// hipExtLaunchKernelGGL, which stamps both from the kernel's own dispatch.
// hipEventRecord around the launch instead puts two marker packets in the
// stream, which cost ~8 us per launch on the 100k-pair DTW step; a timed
// dispatch still costs ~6 us there, hence the sampling stride
// (tsm_profile_enable(n): every n-th launch).  Both events stay null (no
// timing) inside a capture, with profiling off and between samples.
struct ProfEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    ProfEvents() {
        if (!g_ctx.prof || g_ctx.capturing) return;
        if (g_ctx.prof_seen++ % g_ctx.prof_every != 0) return;
        if (g_ctx.ev_used == g_ctx.ev.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess) return;
            if (hipEventCreate(&b) != hipSuccess) {
                (void)hipEventDestroy(a);
                return;
            }
            g_ctx.ev.push_back({a, b});
        }
        start = g_ctx.ev[g_ctx.ev_used].first;
        stop = g_ctx.ev[g_ctx.ev_used].second;
        g_ctx.ev_used++;
    }
};

struct BucketTable;

// DP launcher for one measure and precision; lanes_hint 0 (table), 4 or 8.
// Defined in tsm_dp_<measure>.hip.
template <int M, bool F64>
int run_dp(Workspace &w, int lanes_hint, const tsm_trajset &set, const PairSpace &ps, int64_t npairs, double thr,
           double *out, int32_t *counts, hipStream_t st, double delta = 0.0, double gx = 0.0, double gy = 0.0);

// Argument checks shared by the entry points (tsm_api.hip); prm receives
// the defaults overridden by prm_in.
int validate(int measure, const tsm_trajset *set, int64_t npairs, const tsm_params *prm_in, const double *out,
             tsm_params &prm);

// The kernel launches of one call on workspace w, stream-ordered on st
// (g_mu held).
int launch(Workspace &w, int measure, const tsm_trajset *set, const PairSpace &ps, int64_t npairs,
           const tsm_params &prm, double *out, int32_t *counts, hipStream_t st);

}  // namespace tsm
