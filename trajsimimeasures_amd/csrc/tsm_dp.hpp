// tsm_dp.hpp -- the wavefront dynamic-programming kernel for DTW, EDR, LCSS,
// ERP and discrete Frechet.
//
// Recurrences (1-based DP matrix D of (nrow+1) x (ncol+1)):
//   DTW      D = d + min(D[i-1][j-1], D[i-1][j], D[i][j-1]),  inf boundary   core/dtw.py:27-53
//   Frechet  D = max(d, min3(...)), inf boundary, run on squared distances  core/discret_frechet.py:26-50
//   EDR      D = min(diag + (match ? 0 : 1), up + 1, left + 1), zero boundary core/edr.py:29-52
//   LCSS     D = match ? diag + 1 : max(up, left), zero boundary             core/lcss.py:31-53
//   ERP      D = min(left + gb_j, up + ga_i, diag + d), D[i][0] = Sa,
//            D[0][j] = Sb (constant sums of gap distances)                  core/erp.py:29-91
// All five are symmetric in their two trajectories, so the engine puts the
// longer trajectory on the rows (streamed) and the shorter on the columns
// (held in registers).
//
// Layout of the work: a "segment" of L lanes (L = 4, 8 or 16; 64/L segments
// per wavefront) owns one pair.  Lane j of a segment holds W consecutive
// columns (their points in VGPRs and the previous row of D for them) and
// sweeps the rows in skewed order: at step s it computes row s - j.  The
// value of D at the lane's left edge comes from lane j-1, computed one step
// earlier, through one DPP row_shr:1 move; the diagonal is the previous
// step's left value.  No DP matrix touches memory.  Pairs wider than L*W
// columns are swept in column strips; the last lane of a strip writes its
// right-edge column to a scratch vector which the next strip's first lane
// reads back.
//
// Scheduling: a planner sorts pairs by (W bucket, rows) with a counting sort,
// packs them 64/L to a task (all segments of a task share W), and a
// persistent kernel pulls tasks from an atomic counter, largest first.
#pragma once

#include <type_traits>

#include "tsm_device.hpp"

namespace tsm {

constexpr int kRowKeys = 1024;  // rows are bucketed as min(nrow, kRowKeys - 1) for the sort

// Maximum register columns per lane for a segment width L.
template <int L> struct WMax;
template <> struct WMax<4> { static constexpr int value = 32; };
template <> struct WMax<8> { static constexpr int value = 24; };
template <> struct WMax<16> { static constexpr int value = 16; };

struct DPArgs {
    const float *xy32;
    const double *xy64;
    const int64_t *offs;
    const float *gap32;      // ERP: per-point gap distance, FP32
    const double *gap64;     // ERP: per-point gap distance, FP64
    const float *gsum32;     // ERP: per-trajectory sum of gap distances
    const double *gsum64;
    PairSpace ps;
    const uint32_t *slots;   // task-major pair slots, kEmptySlot = idle segment
    const int64_t *wend;     // wend[k]: end of the task range of bucket W = WMAX - k
    const int64_t *ntasks;   // device scalar
    unsigned int *counter;   // task queue head
    void *bnd;               // strip boundary scratch (AT), per wave and segment
    int64_t bnd_stride;
    double thr;              // EDR/LCSS squared-distance threshold
    double *out;
    int32_t *counts;
};

// Per-measure policy: coordinate type CT, accumulator type AT, boundary
// values, cell update and the final value.
template <int M, bool F64> struct Rec;

template <bool F64> struct RecDTW {
    using CT = typename std::conditional<F64, double, float>::type;
    using AT = CT;
    static constexpr bool kGap = false;
    __device__ static AT top(AT) { return (AT)__builtin_inf(); }
    __device__ static AT left0(AT) { return (AT)__builtin_inf(); }
    __device__ static AT cell(CT ax, CT ay, AT, CT bx, CT by, AT, AT diag, AT up, AT left, double) {
        if constexpr (F64) return sqrt(dist2_f64(ax, ay, bx, by)) + min3d(diag, up, left);
        else return dist_f32(ax, ay, bx, by) + min3f(diag, up, left);
    }
    __device__ static double finish(AT v, int64_t, int64_t, int32_t *) { return (double)v; }
};
template <bool F64> struct Rec<kDTW, F64> : RecDTW<F64> {};

template <bool F64> struct RecFrechet {
    using CT = typename std::conditional<F64, double, float>::type;
    using AT = CT;
    static constexpr bool kGap = false;
    __device__ static AT top(AT) { return (AT)__builtin_inf(); }
    __device__ static AT left0(AT) { return (AT)__builtin_inf(); }
    // Squared domain: sqrt is monotone, so max/min commute with it and one
    // square root of the final cell equals the reference's per-cell sqrt.
    __device__ static AT cell(CT ax, CT ay, AT, CT bx, CT by, AT, AT diag, AT up, AT left, double) {
        if constexpr (F64) return fmax(dist2_f64(ax, ay, bx, by), min3d(diag, up, left));
        else return fmaxf(dist2_f32(ax, ay, bx, by), min3f(diag, up, left));
    }
    __device__ static double finish(AT v, int64_t, int64_t, int32_t *) { return sqrt((double)v); }
};
template <bool F64> struct Rec<kFrechet, F64> : RecFrechet<F64> {};

template <bool F64> struct RecERP {
    using CT = typename std::conditional<F64, double, float>::type;
    using AT = CT;
    static constexpr bool kGap = true;
    __device__ static AT top(AT sb) { return sb; }
    __device__ static AT left0(AT sa) { return sa; }
    __device__ static AT cell(CT ax, CT ay, AT ga, CT bx, CT by, AT gb, AT diag, AT up, AT left, double) {
        if constexpr (F64) return min3d(left + gb, up + ga, diag + sqrt(dist2_f64(ax, ay, bx, by)));
        else return min3f(left + gb, up + ga, diag + dist_f32(ax, ay, bx, by));
    }
    __device__ static double finish(AT v, int64_t, int64_t, int32_t *) { return (double)v; }
};
template <bool F64> struct Rec<kERP, F64> : RecERP<F64> {};

// EDR / LCSS: integer counts carried in FP32 (exact below 2^24); the match
// test uses FP64 coordinates and the FP64 threshold so it reproduces the
// FP64 oracle's decision bit for bit.
template <bool F64> struct Rec<kEDR, F64> {
    using CT = double;
    using AT = float;
    static constexpr bool kGap = false;
    __device__ static AT top(AT) { return 0.0f; }
    __device__ static AT left0(AT) { return 0.0f; }
    __device__ static AT cell(CT ax, CT ay, AT, CT bx, CT by, AT, AT diag, AT up, AT left, double thr) {
        AT sub = dist2_f64(ax, ay, bx, by) < thr ? 0.0f : 1.0f;
        return fminf(diag + sub, fminf(up, left) + 1.0f);
    }
    __device__ static double finish(AT v, int64_t na, int64_t nb, int32_t *cnt) {
        int32_t c = (int32_t)v;
        if (cnt) *cnt = c;
        return (double)c / (double)(na > nb ? na : nb);
    }
};
template <bool F64> struct Rec<kLCSS, F64> {
    using CT = double;
    using AT = float;
    static constexpr bool kGap = false;
    __device__ static AT top(AT) { return 0.0f; }
    __device__ static AT left0(AT) { return 0.0f; }
    __device__ static AT cell(CT ax, CT ay, AT, CT bx, CT by, AT, AT diag, AT up, AT left, double thr) {
        return dist2_f64(ax, ay, bx, by) < thr ? diag + 1.0f : fmaxf(up, left);
    }
    __device__ static double finish(AT v, int64_t na, int64_t nb, int32_t *cnt) {
        int32_t c = (int32_t)v;
        if (cnt) *cnt = c;
        return 1.0 - (double)c / (double)(na < nb ? na : nb);
    }
};

template <typename CT> struct Pt;
template <> struct Pt<float> {
    __device__ static void load(const DPArgs &a, int64_t i, float &x, float &y) {
        float2 v = reinterpret_cast<const float2 *>(a.xy32)[i];
        x = v.x;
        y = v.y;
    }
};
template <> struct Pt<double> {
    __device__ static void load(const DPArgs &a, int64_t i, double &x, double &y) {
        double2 v = reinterpret_cast<const double2 *>(a.xy64)[i];
        x = v.x;
        y = v.y;
    }
};
template <typename AT> __device__ __forceinline__ AT gap_at(const DPArgs &a, int64_t i);
template <> __device__ __forceinline__ float gap_at<float>(const DPArgs &a, int64_t i) { return a.gap32[i]; }
template <> __device__ __forceinline__ double gap_at<double>(const DPArgs &a, int64_t i) { return a.gap64[i]; }
template <typename AT> __device__ __forceinline__ AT gsum_at(const DPArgs &a, int64_t t);
template <> __device__ __forceinline__ float gsum_at<float>(const DPArgs &a, int64_t t) { return a.gsum32[t]; }
template <> __device__ __forceinline__ double gsum_at<double>(const DPArgs &a, int64_t t) { return a.gsum64[t]; }

// One task: 64/L segments, each sweeping one pair with W columns per lane.
template <int M, bool F64, int L, int W>
__device__ __forceinline__ void run_task(const DPArgs &a, int64_t task, int64_t wave_id) {
    using R = Rec<M, F64>;
    using CT = typename R::CT;
    using AT = typename R::AT;
    constexpr int S = kWave / L;
    constexpr int kSpan = L * W;
    const int lane = threadIdx.x & (kWave - 1);
    const int seg = lane / L;
    const int j = lane % L;

    const uint32_t slot = a.slots[task * S + seg];
    const bool valid = slot != kEmptySlot;
    int64_t ia = 0, ib = 0;
    if (valid) decode_pair(a.ps, (int64_t)slot, ia, ib);
    const int64_t oa = a.offs[ia], na = a.offs[ia + 1] - oa;
    const int64_t ob = a.offs[ib], nb = a.offs[ib + 1] - ob;
    const bool swp = na < nb;  // rows = longer trajectory
    const int64_t orow = swp ? ob : oa, ocol = swp ? oa : ob;
    const int nrow = valid ? (int)(swp ? nb : na) : 0;
    const int ncol = valid ? (int)(swp ? na : nb) : 0;
    const int64_t trow = swp ? ib : ia, tcol = swp ? ia : ib;

    AT sa = 0, sb = 0;
    if constexpr (R::kGap) {
        sa = gsum_at<AT>(a, trow);
        sb = gsum_at<AT>(a, tcol);
    }
    const AT topv = R::top(sb);
    const AT leftv = R::left0(sa);

    // a pair needing strips longer than the scratch column (max_len was
    // understated by the caller) is not computed; its result is NaN
    const bool fits = (ncol <= kSpan) || (nrow <= a.bnd_stride);
    if (valid && !fits && j == 0) a.out[slot] = __builtin_nan("");
    const int nstrips = (valid && fits) ? (ncol + kSpan - 1) / kSpan : 0;
    const int maxrow = __builtin_amdgcn_readfirstlane(wave_max(nrow));
    const int maxstrips = __builtin_amdgcn_readfirstlane(wave_max(nstrips));
    const int nsteps = maxrow + L - 1;
    AT *bnd = reinterpret_cast<AT *>(a.bnd) + (wave_id * S + seg) * a.bnd_stride;
    const int rlast = nrow > 0 ? nrow - 1 : 0;

    for (int strip = 0; strip < maxstrips; ++strip) {
        const bool in_strip = strip < nstrips;
        const int col0 = strip * kSpan + j * W;  // first column (0-based) of this lane
        CT bx[W], by[W];
        AT gb[W], prev[W];
#pragma unroll
        for (int c = 0; c < W; ++c) {
            int col = col0 + c;
            col = col < ncol ? col : (ncol > 0 ? ncol - 1 : 0);
            Pt<CT>::load(a, ocol + col, bx[c], by[c]);
            if constexpr (R::kGap) gb[c] = gap_at<AT>(a, ocol + col);
            else gb[c] = 0;
            prev[c] = topv;
        }
        const AT corner0 = (col0 == 0) ? (AT)0 : topv;  // D[0][col0]
        AT left_prev = corner0;
        AT last = 0;
        const bool read_bnd = (j == 0) && strip > 0;
        const bool write_bnd = (j == L - 1) && (strip + 1 < nstrips);

        const uint32_t nrow_s = in_strip ? (uint32_t)nrow : 0u;  // rows this lane sweeps in this strip
        const AT left_fixed = strip == 0 ? leftv : (AT)0;           // D[i][0] for the first lane of strip 0
        // a_r prefetch, one step ahead; indices clamped into the row trajectory
        const int r0 = min(max(-j, 0), rlast);
        CT ax, ay;
        Pt<CT>::load(a, orow + r0, ax, ay);
        AT ga = 0;
        if constexpr (R::kGap) ga = gap_at<AT>(a, orow + r0);
        AT bl = 0;
        if (read_bnd) bl = bnd[r0];

        for (int s = 0; s < nsteps; ++s) {
            const AT recv = shr1(last);
            const int rn = min(max(s + 1 - j, 0), rlast);
            CT nx, ny;
            Pt<CT>::load(a, orow + rn, nx, ny);
            AT nga = 0;
            if constexpr (R::kGap) nga = gap_at<AT>(a, orow + rn);
            AT nbl = 0;
            if (read_bnd) nbl = bnd[rn];

            const int r = s - j;
            if ((uint32_t)r < nrow_s) {
                const AT left_in = (j == 0) ? (read_bnd ? bl : left_fixed) : recv;
                AT diag = left_prev;  // D[r][col0]: corner0 before the first row
                left_prev = left_in;
                AT lf = left_in;
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    const AT up = prev[c];
                    const AT v = R::cell(ax, ay, ga, bx[c], by[c], gb[c], diag, up, lf, a.thr);
                    diag = up;
                    prev[c] = v;
                    lf = v;
                }
                last = lf;
                if (write_bnd) bnd[r] = lf;
            }
            ax = nx;
            ay = ny;
            ga = nga;
            bl = nbl;
        }

        if (in_strip && strip == nstrips - 1) {
            const int cend = ncol - 1 - strip * kSpan;  // last column within this strip
            if (j == cend / W) {
                const int cr = cend % W;
                AT v = prev[0];
#pragma unroll
                for (int c = 1; c < W; ++c) v = (c == cr) ? prev[c] : v;
                int32_t cnt = 0;
                double res = R::finish(v, na, nb, &cnt);
                a.out[slot] = res;
                if ((M == kEDR || M == kLCSS) && a.counts) a.counts[slot] = cnt;
            }
        }
        if (maxstrips > 1) __threadfence();
    }
}

template <int M, bool F64, int L, int W>
__device__ __forceinline__ bool dispatch_w(const DPArgs &a, int64_t task, int64_t wave_id, int w) {
    if constexpr (W == 0) {
        return false;
    } else {
        if (w == W) {
            run_task<M, F64, L, W>(a, task, wave_id);
            return true;
        }
        return dispatch_w<M, F64, L, W - 1>(a, task, wave_id, w);
    }
}

// Persistent kernel: each wavefront pulls tasks until the queue is drained.
template <int M, bool F64, int L>
__global__ __launch_bounds__(kWave) void dp_kernel(DPArgs a) {
    constexpr int WM = WMax<L>::value;
    const int64_t ntasks = *a.ntasks;
    const int64_t wave_id = blockIdx.x;
    const int lane = threadIdx.x;
    while (true) {
        unsigned int t = 0;
        if (lane == 0) t = atomicAdd(a.counter, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if ((int64_t)t >= ntasks) break;
        int w = WM;
#pragma unroll 1
        for (int k = 0; k < WM; ++k) {
            if ((int64_t)t < a.wend[k]) {
                w = WM - k;
                break;
            }
        }
        w = __builtin_amdgcn_readfirstlane(w);
        dispatch_w<M, F64, L, WM>(a, (int64_t)t, wave_id, w);
    }
}

}  // namespace tsm
