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
// Layout of the work: a "segment" of L lanes (L = 2, 4, 8, 16 or 64, so 32,
// 16, 8, 4 or 1 segments per wavefront) owns one pair.  Lane j of a segment holds W
// consecutive columns -- their points in VGPRs and the previous row of D for
// them -- and sweeps the rows in skewed order: at step s it computes row
// s - j.  The value of D at the lane's left edge comes from lane j-1,
// computed one step earlier, through one DPP row_shr:1 move (wave_shr:1 at
// L = 64); the diagonal is
// the previous step's left value.  No DP matrix touches memory.  The row
// state is double-buffered across an unrolled pair of steps (even steps read
// buffer A and write B, odd steps the reverse), so a row update needs no
// register moves.  Pairs wider than L*W columns are swept in column strips;
// the last lane of a strip writes its right-edge column to a scratch vector
// which the next strip's first lane reads back.
//
// Scheduling: every pair falls in a bucket b = (L, W) chosen per column
// count by a table (cost model or measured autotuning); a planner sorts
// pairs by (bucket, rows) with a counting sort and packs them 64/L to a task,
// so all segments of a task share (L, W) and have similar row counts; one
// persistent kernel pulls tasks from an atomic counter, largest W first.
// Small batches skip all of it (direct launches: wave w, pair w).
#pragma once

#include <type_traits>

#include "tsm_device.hpp"

namespace tsm {


// Row bins per bucket: bin = min(nrow, 511) * kRowKeys / 512.
#ifndef TSM_ROW_KEYS
#define TSM_ROW_KEYS 128
#endif
constexpr int kRowKeys = TSM_ROW_KEYS;
static_assert(kRowKeys >= 16 && kRowKeys <= 128 && (kRowKeys & (kRowKeys - 1)) == 0, "row keys: power of two <= 128");
constexpr int kMaxSeg = 16;     // segments per wave at L = 4 (the strip scratch; 2-lane pairs never need strips)
constexpr int kMaxBuckets = 96;  // the planner's bucket capacity
constexpr int kLongRowCols16 = 128;  // long-row pairs: 16 lanes up to this many columns, 64 beyond
constexpr int kTabMax = 257;    // bucket table covers column counts 1..256

// Register columns per lane: 24 with FP32 coordinates (TSM_WMAX32; measured
// faster than 32, whose ~190 VGPRs halve occupancy) and 24 with FP64 ones.
#ifndef TSM_WMAX32
#define TSM_WMAX32 24
#endif
template <typename CT> struct WMaxOf { static constexpr int value = sizeof(CT) == 4 ? TSM_WMAX32 : 24; };

// Buckets (L lanes per pair, W register columns per lane), with N64 = WM / 4
// 64-lane buckets:
//   b < N64: L = 64, W = WM - 4 b (24, 20, ..., 4 at WM = 24).  A pair on a
//     whole wavefront takes a quarter of the steps of 16 lanes and the
//     smallest W that holds it the fewest cells per step, which is what
//     counts wherever a pair's own chain of steps, not the total work, sets
//     the time: pairs wider than 16 * WM columns in launches of at most
//     2^20 pairs (W the smallest with 64 W >= its columns, W = WM beyond
//     64 * WM in strips of 64 * WM columns), and every pair of a direct
//     launch (few pairs, one wavefront each, no planner);
//   N64 <= b < N64 + WM: L = 16, W = WM - (b - N64) (b = N64 also takes the
//     pairs wider than the table otherwise, in strips of 16 * WM columns);
//   N64 + WM <= b < N64 + 3 WM: b' = b - N64 - WM, L = (b' odd ? 4 : 8),
//     W = WM - b' / 2;
//   N64 + 3 WM <= b: L = 2, W = WM / 2 - (b - N64 - 3 WM) (short columns).
// Tasks run in bucket order, so the widest pairs start first.
__host__ __device__ constexpr int bucket_n64(int WM) { return WM / 4; }
__host__ __device__ constexpr int bucket_count(int WM) { return WM / 4 + 3 * WM + WM / 2; }
__host__ __device__ constexpr int bucket_of(int WM, int L, int W) {
    return L == 64   ? (WM - W) / 4
           : L == 16 ? WM / 4 + WM - W
           : L == 2  ? WM / 4 + 3 * WM + (WM / 2 - W)
                     : WM / 4 + WM + 2 * (WM - W) + (L == 4 ? 1 : 0);
}
__host__ __device__ constexpr int bucket_L(int WM, int b) {
    return b < WM / 4 ? 64 : b < WM / 4 + WM ? 16 : b < WM / 4 + 3 * WM ? (((b - WM / 4 - WM) & 1) ? 4 : 8) : 2;
}
__host__ __device__ constexpr int bucket_W(int WM, int b) {
    return b < WM / 4            ? WM - 4 * b
           : b < WM / 4 + WM     ? WM - (b - WM / 4)
           : b < WM / 4 + 3 * WM ? WM - (b - WM / 4 - WM) / 2
                                 : WM / 2 - (b - WM / 4 - 3 * WM);
}
// The 64-lane bucket of a pair with nc columns: the smallest W (a multiple
// of 4) with 64 W >= nc, W = WM beyond 64 * WM (strips).
__host__ __device__ constexpr int bucket_w64(int WM, int64_t nc) {
    return nc > 64 * (int64_t)WM ? 0 : (WM - 4 * (int)((nc + 255) / 256)) / 4;
}
// Largest W of an L-lane bucket.
__host__ __device__ constexpr int bucket_wmax(int WM, int L) { return L == 2 ? WM / 2 : WM; }

// Column count -> bucket table for one launch.  The planner and the task
// queue work on the launch's used buckets only, numbered 0 .. nb - 1 in
// schedule order (widest W first: the longest tasks start first); kid maps
// that number to the (L, W) kernel variant above.
struct BucketTable {
    uint8_t b[kTabMax];  // used-bucket number of a pair whose shorter side has ncol columns, 1 <= ncol <= ncmax
    uint8_t lanes[kMaxBuckets];  // L of used bucket k
    uint8_t kid[kMaxBuckets];    // (L, W) variant of used bucket k (bucket_L / bucket_W)
    uint8_t w64[8];      // with wide64: used-bucket number of L = 64, W = WM - 4 i (i < WM / 4)
    uint8_t w16[32];     // with long_rows: used-bucket number of L = 16, W = w (1 <= w <= kLongRowCols16 / 16)
    int long_rows;       // nonzero: a pair with more rows and at most wide16 columns takes 16 lanes up to
                         // kLongRowCols16 columns and 64 beyond (w64[bucket_w64(wm, ncol)])
    int nb;              // used buckets
    int ncmax;           // largest column count the table covers
    int wide;            // used bucket for ncol > ncmax (16 lanes, W = WM; strips when wider than wide16)
    int wide16;          // with wide64: the column count up to which `wide` still applies (16 * WM)
    int wm;              // with wide64: WM; a wider pair takes w64[bucket_w64(WM, ncol)]
    int wide64;          // nonzero: pairs wider than wide16 take 64-lane segments
};

struct DPArgs {
    const float *xy32;
    const double *xy64;
    const int64_t *offs;
    double gx, gy;           // ERP: the gap point g
    PairSpace ps;
    const uint32_t *slots;   // pair slots grouped by task, kEmptySlot = idle segment; null: a direct
                             // launch (wave w computes pair w on 64 lanes, no planner ran)
    int64_t nset;            // direct launch: trajectories in the set (pair indices outside it give NaN)
    const int64_t *meta;     // [0] ntasks, [1 + b] task end of used bucket b, [1 + nb + b] its slot base
    int nb;                  // number of used buckets
    uint8_t kid[kMaxBuckets];  // (L, W) variant of used bucket b
    unsigned int *counter;   // task queue head
    void *bnd;               // strip boundary scratch (AT), kMaxSeg segments per wave
    int64_t bnd_stride;
    double thr;              // EDR/LCSS squared-distance threshold
    float thr_lo, thr_hi;    // EDR/LCSS FP32 guard band, 0 <= lo <= hi: below lo a match, at or above hi none
    const double *t64;       // STEDR/STLCSS: per-point timestamps
    double delta;            // STEDR/STLCSS: time threshold
    double *out;
    int32_t *counts;
    uint32_t *plan_zero;     // the planner's histogram and fill counters, zeroed here for the next call
    int plan_zero_n;         // their words
};

// Per-measure policy: coordinate type CT, accumulator type AT, boundary
// values, cell update and the final value.
template <int M, bool F64> struct Rec;

template <bool F64> struct RecDTW {
    using CT = typename std::conditional<F64, double, float>::type;
    using AT = CT;
    static constexpr int kWMax = WMaxOf<CT>::value;
    static constexpr bool kTime = false;
    static constexpr bool kBand = false;
    static constexpr bool kGap = false;
    static constexpr bool kIdxBnd = false;  // boundaries are top() / left0()
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
    static constexpr int kWMax = WMaxOf<CT>::value;
    static constexpr bool kTime = false;
    static constexpr bool kBand = false;
    static constexpr bool kGap = false;
    static constexpr bool kIdxBnd = false;  // boundaries are top() / left0()
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
    static constexpr int kWMax = WMaxOf<CT>::value;
    static constexpr bool kTime = false;
    static constexpr bool kBand = false;
    static constexpr bool kGap = true;
    static constexpr bool kIdxBnd = false;  // boundaries are top() / left0()
    __device__ static AT top(AT sb) { return sb; }
    __device__ static AT left0(AT sa) { return sa; }
    __device__ static AT cell(CT ax, CT ay, AT ga, CT bx, CT by, AT gb, AT diag, AT up, AT left, double) {
        if constexpr (F64) return min3d(left + gb, up + ga, diag + sqrt(dist2_f64(ax, ay, bx, by)));
        else return min3f(left + gb, up + ga, diag + dist_f32(ax, ay, bx, by));
    }
    __device__ static double finish(AT v, int64_t, int64_t, int32_t *) { return (double)v; }
};
template <bool F64> struct Rec<kERP, F64> : RecERP<F64> {};

// EDR / LCSS counts.  The match decision must equal the FP64 oracle's,
// (dx*dx + dy*dy) < thr on FP64 coordinates, bit for bit.  Two forms:
//   F64 = true (and STEDR/STLCSS): the test itself runs on the FP64
//     coordinates; counts are carried in FP32 (exact below 2^24) and the
//     update selects with the compare's lane mask:
//       EDR  match ? diag : min3(diag, up, left) + 1 (unit-cost edit
//            distances satisfy diag <= min(up, left) + 1),
//       LCSS match ? diag + 1 : max(up, left).
//   F64 = false: the test runs on the FP32 copy against a guard band
//     [thr_lo, thr_hi) that provably contains every FP32 squared distance
//     whose FP64 counterpart could fall on the other side of thr (the host
//     derives it from a bound on |coordinate|); a lane that meets a value
//     inside the band recomputes its whole row step with the FP64 test.
//     Counts are int32 and the match is a bit, not a lane mask: with
//     0 <= thr_lo <= thr_hi (as the host guarantees) and d2 >= 0, the FP32
//     bit patterns order as the values, so t = bits(d2) - bits(thr_lo) has
//     its sign bit set exactly when d2 < thr_lo, and one unsigned compare
//     t < bits(thr_hi) - bits(thr_lo) finds the band.  A compare writes a
//     scalar mask and the chip issues fewer of those than plain VALU
//     operations (a cell-chain test: 5.35e12 cells/s with the bit against
//     4.77e12 with two compares and a select, 4.39e12 with a carry-in add);
//     the update is then an add and one v_max3_i32:
//       EDR keeps N[i][j] = (i + j) - E[i][j] instead of the edit distance
//         E.  Substituting E = (i + j) - N into E = min(diag + 1 - match,
//         up + 1, left + 1) gives N = max(N_diag + 1 + match, N_up,
//         N_left), exactly.  The reference's E is zero on both boundaries,
//         so N[0][j] = j and N[i][0] = i; the count is na + nb - N at the
//         last cell.
//       LCSS: match ? diag + 1 : max(up, left) equals max(up, left, diag +
//         match), because LCSS values of prefixes never decrease and grow
//         by at most one per added point (diag <= up, left <= diag + 1).
template <int M, typename AT_> struct RecCount {
    using AT = AT_;
    static constexpr bool kInt = std::is_same<AT, int32_t>::value;
    static constexpr bool kGap = false;
    static constexpr bool kEdit = M == kEDR || M == kSTEDR;  // EDR form, else LCSS form
    // the int EDR form's N has the boundaries N[0][j] = j, N[i][0] = i
    // (core/edr.py:30-32 zeroes E's); the others keep zero boundaries
    static constexpr bool kIdxBnd = kEdit && kInt;
    __device__ static AT top(AT) { return 0; }
    __device__ static AT left0(AT) { return 0; }
    // int forms, m in {0, 1}
    __device__ static AT pick_bit(int32_t m, AT diag, AT up, AT left) {
        if constexpr (kEdit) return max(max(diag + 1 + m, up), left);
        else return max(max(diag + m, up), left);
    }
    __device__ static AT pick(bool match, AT diag, AT up, AT left) {
        if constexpr (kInt) return pick_bit((int32_t)match, diag, up, left);
        else if constexpr (kEdit) return match ? diag : min3f(diag, up, left) + 1.0f;
        else return match ? diag + 1.0f : fmaxf(up, left);
    }
    __device__ static double finish(AT v, int64_t na, int64_t nb, int32_t *cnt) {
        const int32_t c = kIdxBnd ? (int32_t)(na + nb) - (int32_t)v : (int32_t)v;
        if (cnt) *cnt = c;
        if constexpr (kEdit) return (double)c / (double)(na > nb ? na : nb);
        else return 1.0 - (double)c / (double)(na < nb ? na : nb);
    }
};
template <int M> struct RecCount64 : RecCount<M, float> {
    using CT = double;
    static constexpr int kWMax = WMaxOf<CT>::value;
    static constexpr bool kTime = false;
    static constexpr bool kBand = false;
    __device__ static float cell(CT ax, CT ay, float, CT bx, CT by, float, float diag, float up, float left,
                                 double thr) {
        return RecCount<M, float>::pick(dist2_f64(ax, ay, bx, by) < thr, diag, up, left);
    }
};
template <int M> struct RecCount32 : RecCount<M, int32_t> {
    using CT = float;
    static constexpr int kWMax = WMaxOf<CT>::value;
    static constexpr bool kTime = false;
    static constexpr bool kBand = true;
};
// STEDR / STLCSS (core/stedr.py:44-46, core/stlcss.py:41-42): a match also
// needs |t1 - t2| < delta, tested in FP64 like the spatial part.  Columns
// carry a third FP64 value, so at most 16 per lane keep the registers.
// (STEDR with int counts and the match as the sign bits of d2 - thr and
// |t1 - t2| - delta measured slower in this kernel, 0.567 -> 0.647 ms per
// 100k pairs, though faster in a bare cell chain.)
template <int M> struct RecST : RecCount<M, float> {
    using CT = double;
    static constexpr int kWMax = 16;
    static constexpr bool kTime = true;
    static constexpr bool kBand = false;
    __device__ static float cell_t(CT ax, CT ay, CT at, CT bx, CT by, CT bt, float diag, float up, float left,
                                   double thr, double delta) {
        const bool m = (dist2_f64(ax, ay, bx, by) < thr) && (fabs(at - bt) < delta);
        return RecCount<M, float>::pick(m, diag, up, left);
    }
};
template <> struct Rec<kSTEDR, true> : RecST<kSTEDR> {};
template <> struct Rec<kSTLCSS, true> : RecST<kSTLCSS> {};
template <> struct Rec<kEDR, true> : RecCount64<kEDR> {};
template <> struct Rec<kLCSS, true> : RecCount64<kLCSS> {};
template <> struct Rec<kEDR, false> : RecCount32<kEDR> {};
template <> struct Rec<kLCSS, false> : RecCount32<kLCSS> {};

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
// ERP gap distance |p - g| (core/erp.py:32-40), computed where the point is
// used instead of in a pass over the whole set.  FP32: p is a point of the
// FP32 copy and g its position in that copy's frame (gap_point_f32); FP64:
// the oracle's expression on the FP64 coordinates.
__device__ __forceinline__ float gap_of(float x, float y, float gx, float gy) { return dist_f32(x, y, gx, gy); }
__device__ __forceinline__ double gap_of(double x, double y, double gx, double gy) {
    return sqrt(dist2_f64(x, y, gx, gy));
}

// The gap point in the frame of the coordinates the kernel reads.  The FP32
// copy is xy64 minus an origin the kernel is not told, so g is placed
// relative to a point r of the set: g32 = r32 - (float)(r64 - g), with the
// FP64 subtraction exact to rounding; |p32 - g32| then differs from
// |p - g| by FP32 roundings of the set's extent, as the cell distances do.
template <typename CT> __device__ __forceinline__ void gap_point(const DPArgs &a, int64_t r, CT &gx, CT &gy);
template <> __device__ __forceinline__ void gap_point<float>(const DPArgs &a, int64_t r, float &gx, float &gy) {
    const float2 r32 = reinterpret_cast<const float2 *>(a.xy32)[r];
    const double2 r64 = reinterpret_cast<const double2 *>(a.xy64)[r];
    gx = r32.x - (float)(r64.x - a.gx);
    gy = r32.y - (float)(r64.y - a.gy);
}
template <> __device__ __forceinline__ void gap_point<double>(const DPArgs &a, int64_t, double &gx, double &gy) {
    gx = a.gx;
    gy = a.gy;
}

// Sum of the gap distances of the n points from o (core/erp.py:44-49): the
// L lanes of a segment take every L-th point, accumulate in FP64 and
// reduce across the segment (its lanes are an aligned group of L).
template <typename CT, int L>
__device__ __forceinline__ double gap_sum(const DPArgs &a, int64_t o, int n, int j, CT gx, CT gy) {
    double s = 0.0;
    for (int i = j; i < n; i += L) {
        CT x, y;
        Pt<CT>::load(a, o + i, x, y);
        s += (double)gap_of(x, y, gx, gy);
    }
#pragma unroll
    for (int m = 1; m < L; m <<= 1) s += __shfl_xor(s, m, kWave);
    return s;
}

// Lane shift by one: within rows of 16 lanes for segments of up to 16
// lanes, across the wavefront for 64; sources outside read 0 (every
// segment's lane 0 replaces the value).
template <int L> __device__ __forceinline__ float shl_in(float v) {
    if constexpr (L > 16) return shr1_wave(v);
    else return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x111, 0xF, 0xF, true));
}
template <int L> __device__ __forceinline__ int32_t shl_in(int32_t v) {
    return __builtin_bit_cast(int32_t, shl_in<L>(__builtin_bit_cast(float, v)));
}
template <int L> __device__ __forceinline__ double shl_in(double v) {
    if constexpr (L > 16) return shr1_wave(v);
    else return shr1(v);
}

// Per-lane state of one strip sweep.
template <typename CT, typename AT> struct RowPt {
    CT x, y;
    CT t;  // timestamp (STEDR/STLCSS)
};

// One step of the sweep: row r = s - j of this lane, reading the previous
// row from `src` and writing the new one to `dst`.  `cur` holds the point of
// row r, loaded one step earlier; `nxt` receives the point of the lane's
// next row.  (Loading two steps ahead measured slower.)
template <int M, bool F64, int L, int W, bool kBnd, bool kLds>
__device__ __forceinline__ void dp_step(const DPArgs &a, int s, int j, uint32_t nrow_s, int rlast, int64_t orow,
                                        int64_t ocol, int col0, int ncol, bool lane0, typename Rec<M, F64>::AT left_fixed,
                                        typename Rec<M, F64>::AT *bnd, bool write_bnd,
                                        const typename Rec<M, F64>::CT (&bx)[W],
                                        const typename Rec<M, F64>::CT (&by)[W],
                                        const typename Rec<M, F64>::AT (&gb)[W],
                                        const typename Rec<M, F64>::CT (&bt)[W],
                                        const typename Rec<M, F64>::AT (&src)[W], typename Rec<M, F64>::AT (&dst)[W],
                                        const RowPt<typename Rec<M, F64>::CT, typename Rec<M, F64>::AT> &cur,
                                        RowPt<typename Rec<M, F64>::CT, typename Rec<M, F64>::AT> &nxt,
                                        typename Rec<M, F64>::AT &lp_in, typename Rec<M, F64>::AT &lp_out,
                                        typename Rec<M, F64>::AT cur_bl, typename Rec<M, F64>::AT &nxt_bl,
                                        typename Rec<M, F64>::CT gx, typename Rec<M, F64>::CT gy,
                                        const typename Rec<M, F64>::CT *srow, const double *strow) {
    using R = Rec<M, F64>;
    using CT = typename R::CT;
    using AT = typename R::AT;
    const AT recv = shl_in<L>(src[W - 1]);  // lane j-1's last column of row s - j
    // next row of this lane, kept inside the trajectory: a lane whose next
    // row is out of range does not use the point, so any row will do
    const uint32_t rn = min((uint32_t)(s + 1 - j), (uint32_t)rlast);
    if constexpr (kLds) {  // the row trajectory staged in LDS
        nxt.x = srow[2 * rn];
        nxt.y = srow[2 * rn + 1];
        if constexpr (R::kTime) nxt.t = strow[rn];
    } else {
        Pt<CT>::load(a, orow + rn, nxt.x, nxt.y);
        if constexpr (R::kTime) nxt.t = a.t64[orow + rn];
    }
    if constexpr (kBnd) {
        if (lane0) nxt_bl = bnd[rn];
    }
    const int r = s - j;
    if ((uint32_t)r < nrow_s) {
        AT left_in;
        if constexpr (kBnd) left_in = lane0 ? cur_bl : recv;
        else if constexpr (R::kIdxBnd) left_in = lane0 ? (AT)(r + 1) : recv;  // D[r + 1][0] = r + 1
        else left_in = lane0 ? left_fixed : recv;
        AT diag = lp_in;  // D[r][col0]: the corner value before the first row
        lp_out = left_in;
        AT ga = 0;  // ERP: the row point's gap distance
        if constexpr (R::kGap) ga = (AT)gap_of(cur.x, cur.y, gx, gy);
        AT lf = left_in;
        if constexpr (R::kBand) {
            // FP32 match test; `amb` collects the lanes that met a squared
            // distance inside the guard band (wave-uniform mask)
            uint64_t amb = 0;
            const uint32_t lob = __builtin_bit_cast(uint32_t, a.thr_lo);
            const uint32_t span = __builtin_bit_cast(uint32_t, a.thr_hi) - lob;
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const float d2 = dist2_f32(cur.x, cur.y, bx[c], by[c]);
                const uint32_t t = __builtin_bit_cast(uint32_t, d2) - lob;  // sign bit: d2 < thr_lo
                amb |= __builtin_amdgcn_ballot_w64(t < span);                // thr_lo <= d2 < thr_hi
                const AT v = R::pick_bit((int32_t)(t >> 31), diag, src[c], lf);
                diag = src[c];
                dst[c] = v;
                lf = v;
            }
            if (amb != 0) {
                // the empty asm keeps this lane's test (three VALU
                // instructions) inside the rarely taken branch; hoisted above
                // it, every row step paid for it
                uint64_t am = amb;
                asm volatile("" : "+s"(am));
                if ((am >> (threadIdx.x & (kWave - 1))) & 1ull) {
                    // redo this lane's row step with the FP64 test; src (the
                    // previous row), the corner and the left input are intact
                    const double2 pr = reinterpret_cast<const double2 *>(a.xy64)[orow + r];
                    // the column addresses are loop-invariant; the empty asm
                    // keeps the compiler from hoisting them (and their
                    // registers) out of this rarely taken branch.  Columns
                    // past ncol are padding whose values nothing reads.
                    const double2 *pc = reinterpret_cast<const double2 *>(a.xy64) + ocol + col0;
                    asm volatile("" : "+v"(pc));
                    const int nvalid = ncol - col0;
                    AT dg = lp_in, lx = left_in;
#pragma unroll
                    for (int c = 0; c < W; ++c) {
                        if (c < nvalid) {
                            const double2 q = pc[c];
                            // an opaque copy of the previous row's value:
                            // min(diag, up) here must not be shared with the
                            // FP32 pass above, where sharing it splits every
                            // cell's v_min3 into two v_min (one VALU more)
                            AT up = src[c];
                            asm volatile("" : "+v"(up));
                            const AT v = R::pick(dist2_f64(pr.x, pr.y, q.x, q.y) < a.thr, dg, up, lx);
                            dst[c] = v;
                            lx = v;
                        }
                        dg = src[c];
                        // one column load in flight at a time: all W at once
                        // would need 4W more registers in every path
                        asm volatile("" ::: "memory");
                    }
                    lf = lx;
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < W; ++c) {
                AT v;
                if constexpr (R::kTime)
                    v = R::cell_t(cur.x, cur.y, cur.t, bx[c], by[c], bt[c], diag, src[c], lf, a.thr, a.delta);
                else
                    v = R::cell(cur.x, cur.y, ga, bx[c], by[c], gb[c], diag, src[c], lf, a.thr);
                diag = src[c];
                dst[c] = v;
                lf = v;
            }
        }
        if (write_bnd) bnd[r] = lf;
    }
}

// One task: 64/L segments, each sweeping one pair with W columns per lane.
// Dynamic LDS of direct launches: the row trajectory's points (and
// timestamps) of the wave's pair.
extern __shared__ __attribute__((aligned(16))) unsigned char tsm_dp_lds[];

// Bytes of dynamic LDS a direct launch of measure (M, F64) passes: up to
// 64 WM rows, the longest trajectory a direct launch admits.
template <int M, bool F64> constexpr size_t direct_lds_bytes() {
    using R = Rec<M, F64>;
    return (size_t)64 * R::kWMax * (2 * sizeof(typename R::CT) + (R::kTime ? sizeof(double) : 0));
}

// kDirect: a direct launch's task (L = 64, the wave's own pair, no column
// strips): the wave first copies its row trajectory into LDS, so each row
// step reads its point there instead of waiting on a global load that, with
// at most one wave per SIMD, nothing else would cover.
template <int M, bool F64, int L, int W, bool kDirect = false>
__device__ __forceinline__ void run_task(const DPArgs &a, int64_t slot0, int64_t wave_id) {
    using R = Rec<M, F64>;
    using CT = typename R::CT;
    using AT = typename R::AT;
    constexpr int kSpan = L * W;
    const int lane = threadIdx.x & (kWave - 1);
    const int seg = lane / L;
    const int j = lane % L;
    const bool lane0 = j == 0;

    // a direct launch passes the pair itself (one segment per wave)
    const uint32_t slot = a.slots ? a.slots[slot0 + seg] : (uint32_t)slot0;
    const bool valid = slot != kEmptySlot;
    int64_t ia = 0, ib = 0;
    if (valid) decode_pair(a.ps, (int64_t)slot, ia, ib);
    const int64_t oa = a.offs[ia], na = a.offs[ia + 1] - oa;
    const int64_t ob = a.offs[ib], nb = a.offs[ib + 1] - ob;
    const bool swp = na < nb;  // rows = longer trajectory
    const int64_t orow = swp ? ob : oa, ocol = swp ? oa : ob;
    const int nrow = valid ? (int)(swp ? nb : na) : 0;
    const int ncol = valid ? (int)(swp ? na : nb) : 0;

    AT sa = 0, sb = 0;
    CT gx = 0, gy = 0;
    if constexpr (R::kGap) {
        gap_point<CT>(a, orow, gx, gy);
        sa = (AT)gap_sum<CT, L>(a, orow, nrow, j, gx, gy);
        sb = (AT)gap_sum<CT, L>(a, ocol, ncol, j, gx, gy);
    }
    const AT topv = R::top(sb);
    const AT leftv = R::left0(sa);

    CT *srow = nullptr;
    double *strow = nullptr;
    if constexpr (kDirect) {
        constexpr int kRows = 64 * R::kWMax;  // the host admits rows <= 64 WM
        srow = reinterpret_cast<CT *>(tsm_dp_lds);
        strow = reinterpret_cast<double *>(tsm_dp_lds + 2 * kRows * sizeof(CT));
        const int nst = nrow < kRows ? nrow : kRows;
        for (int i = lane; i < nst; i += kWave) {
            CT x, y;
            Pt<CT>::load(a, orow + i, x, y);
            srow[2 * i] = x;
            srow[2 * i + 1] = y;
            if constexpr (R::kTime) strow[i] = a.t64[orow + i];
        }
        __syncthreads();
    }

    // a pair needing strips longer than the scratch column (max_len was
    // understated by the caller) is not computed; its result is NaN and its
    // count -1
    const bool fits = (ncol <= kSpan) || (L >= 4 && nrow <= a.bnd_stride);
    if (valid && !fits && lane0) {
        a.out[slot] = __builtin_bit_cast(double, 0x7ff8000000000000ull);  // quiet NaN
        if ((M == kEDR || M == kLCSS || M == kSTEDR || M == kSTLCSS) && a.counts) a.counts[slot] = -1;
    }
    const int nstrips = (valid && fits) ? (ncol + kSpan - 1) / kSpan : 0;
    const int maxstrips = __builtin_amdgcn_readfirstlane(wave_max(nstrips));
    // (2-lane segments hold at most WM columns: never strips, no scratch)
    AT *bnd = L >= 4 ? reinterpret_cast<AT *>(a.bnd) + (wave_id * kMaxSeg + seg) * a.bnd_stride : nullptr;
    const int rlast = nrow > 0 ? nrow - 1 : 0;

    for (int strip = 0; strip < maxstrips; ++strip) {
        const bool in_strip = strip < nstrips;
        const int col0 = strip * kSpan + j * W;  // first column (0-based) of this lane
        // a segment's lanes past its last column in this strip compute
        // nothing that is read: the sweep ends when the last lane holding a
        // column has done its last row
        const int used = in_strip ? min(L, (ncol - strip * kSpan + W - 1) / W) : 0;
        const int nsteps = __builtin_amdgcn_readfirstlane(wave_max(in_strip ? nrow + used - 1 : 0));
        CT bx[W], by[W], bt[W];
        AT gb[W], A[W], B[W];
#pragma unroll
        for (int c = 0; c < W; ++c) {
            int col = col0 + c;
            col = col < ncol ? col : (ncol > 0 ? ncol - 1 : 0);
            Pt<CT>::load(a, ocol + col, bx[c], by[c]);
            if constexpr (R::kGap) gb[c] = (AT)gap_of(bx[c], by[c], gx, gy);
            else gb[c] = 0;
            if constexpr (R::kTime) bt[c] = a.t64[ocol + col];
            else bt[c] = 0;
            A[c] = R::kIdxBnd ? (AT)(col0 + c + 1) : topv;  // D[0][col0 + c + 1]
            B[c] = A[c];
        }
        AT lpA = R::kIdxBnd ? (AT)col0 : (col0 == 0) ? (AT)0 : topv;  // D[0][col0]
        AT lpB = lpA;
        const bool write_bnd = (j == L - 1) && (strip + 1 < nstrips);
        const uint32_t nrow_s = in_strip ? (uint32_t)nrow : 0u;
        RowPt<CT, AT> p0, p1;  // points of the lane's rows at even / odd steps
        const uint32_t r0 = min((uint32_t)(-j), (uint32_t)rlast);
        p0.t = 0;
        p1.t = 0;
        if constexpr (kDirect) {
            p0.x = srow[2 * r0];
            p0.y = srow[2 * r0 + 1];
            if constexpr (R::kTime) p0.t = strow[r0];
        } else {
            Pt<CT>::load(a, orow + r0, p0.x, p0.y);
            if constexpr (R::kTime) p0.t = a.t64[orow + r0];
        }
        AT bl0 = 0, bl1 = 0;
        if (kDirect || strip == 0) {
            for (int s = 0; s < nsteps; s += 2) {
                dp_step<M, F64, L, W, false, kDirect>(a, s, j, nrow_s, rlast, orow, ocol, col0, ncol, lane0, leftv, bnd, write_bnd, bx, by, gb,
                                                   bt, A, B, p0, p1, lpA, lpB, bl0, bl1, gx, gy, srow, strow);
                if (s + 1 < nsteps)
                    dp_step<M, F64, L, W, false, kDirect>(a, s + 1, j, nrow_s, rlast, orow, ocol, col0, ncol, lane0, leftv, bnd, write_bnd, bx,
                                                       by, gb, bt, B, A, p1, p0, lpB, lpA, bl1, bl0, gx, gy, srow, strow);
            }
        } else if constexpr (!kDirect) {
            if (lane0) bl0 = bnd[r0];
            for (int s = 0; s < nsteps; s += 2) {
                dp_step<M, F64, L, W, true, false>(a, s, j, nrow_s, rlast, orow, ocol, col0, ncol, lane0, leftv, bnd, write_bnd, bx, by, gb, bt,
                                                A, B, p0, p1, lpA, lpB, bl0, bl1, gx, gy, srow, strow);
                if (s + 1 < nsteps)
                    dp_step<M, F64, L, W, true, false>(a, s + 1, j, nrow_s, rlast, orow, ocol, col0, ncol, lane0, leftv, bnd, write_bnd, bx, by,
                                                    gb, bt, B, A, p1, p0, lpB, lpA, bl1, bl0, gx, gy, srow, strow);
            }
        }

        if (in_strip && strip == nstrips - 1) {
            const int cend = ncol - 1 - strip * kSpan;  // last column within this strip
            if (j == cend / W) {
                const int cr = cend % W;
                // the lane's last row (nrow - 1) ran at step nrow - 1 + j:
                // even steps write B, odd steps write A
                const bool inB = ((nrow - 1 + j) & 1) == 0;
                AT v = inB ? B[0] : A[0];
#pragma unroll
                for (int c = 1; c < W; ++c) v = (c == cr) ? (inB ? B[c] : A[c]) : v;
                int32_t cnt = 0;
                double res = R::finish(v, na, nb, &cnt);
                a.out[slot] = res;
                if ((M == kEDR || M == kLCSS || M == kSTEDR || M == kSTLCSS) && a.counts) a.counts[slot] = cnt;
            }
        }
        if (maxstrips > 1) __threadfence();
    }
}

template <int M, bool F64, int WM, int b>
__device__ __forceinline__ void dispatch_bucket(const DPArgs &a, int bk, int64_t slot0, int64_t wave_id) {
    if constexpr (b < bucket_count(WM)) {
        if (bk == b) {
            run_task<M, F64, bucket_L(WM, b), bucket_W(WM, b)>(a, slot0, wave_id);
            return;
        }
        dispatch_bucket<M, F64, WM, b + 1>(a, bk, slot0, wave_id);
    }
}

// A direct launch's 64-lane variants (b < WM / 4), with the row trajectory
// in LDS.
template <int M, bool F64, int WM, int b>
__device__ __forceinline__ void dispatch_direct(const DPArgs &a, int bk, int64_t wave_id) {
    if constexpr (b < bucket_n64(WM)) {
        if (bk == b) {
            run_task<M, F64, 64, bucket_W(WM, b), true>(a, wave_id, wave_id);
            return;
        }
        dispatch_direct<M, F64, WM, b + 1>(a, bk, wave_id);
    }
}

// Waves of the DP kernel that clear the planner's counters first.
constexpr int kPlanZeroWaves = 16;

// Persistent kernel: each wavefront pulls tasks until the queue is drained.
template <int M, bool F64> __device__ __forceinline__ void dp_body(const DPArgs &a) {
    constexpr int WM = Rec<M, F64>::kWMax;
    const int64_t wave_id = blockIdx.x;
    const int lane = threadIdx.x;
    // the planner that ran before this kernel left its histogram and fill
    // counters set; the next call's planner expects them zero
    const int nz = gridDim.x < kPlanZeroWaves ? (int)gridDim.x : kPlanZeroWaves;
    if ((int)blockIdx.x < nz)
        for (int i = blockIdx.x * kWave + lane; i < a.plan_zero_n; i += nz * kWave) a.plan_zero[i] = 0u;
    const bool direct = a.slots == nullptr;
    bool first = true;
    // one call site of the bucket dispatch (every (L, W) variant is inlined
    // there) for persistent waves, which loop until the task queue is
    // drained; a direct launch's wave takes its 64-lane variant (row points
    // in LDS) once
    while (true) {
        int kid;
        int64_t slot0;
        if (direct) {
            // this wave's pair on all 64 lanes, in the 64-lane bucket of its
            // column count (the host chose this mode only for sets whose
            // pairs need no column strips)
            if (!first) break;
            first = false;
            int64_t ia = 0, ib = 0;
            decode_pair(a.ps, wave_id, ia, ib);
            if (ia < 0 || ia >= a.nset || ib < 0 || ib >= a.nset) {
                if (lane == 0) {
                    a.out[wave_id] = __builtin_bit_cast(double, 0x7ff8000000000000ull);  // quiet NaN
                    if ((M == kEDR || M == kLCSS || M == kSTEDR || M == kSTLCSS) && a.counts) a.counts[wave_id] = -1;
                }
                break;
            }
            const int64_t na = a.offs[ia + 1] - a.offs[ia], nb = a.offs[ib + 1] - a.offs[ib];
            const int64_t nc = na < nb ? na : nb;
            kid = __builtin_amdgcn_readfirstlane(bucket_w64(WM, nc < 1 ? 1 : nc));
            dispatch_direct<M, F64, WM, 0>(a, kid, wave_id);
            break;
        } else {
            unsigned int t = 0;
            if (lane == 0) t = atomicAdd(a.counter, 1u);
            t = __builtin_amdgcn_readfirstlane(t);
            if ((int64_t)t >= a.meta[0]) break;
            // bucket of task t: first b with task_end[b] > t (binary search)
            int lo = 0, hi = a.nb - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (a.meta[1 + mid] > (int64_t)t) hi = mid;
                else lo = mid + 1;
            }
            const int bk = __builtin_amdgcn_readfirstlane(lo);
            kid = a.kid[bk];
            const int64_t tstart = bk > 0 ? a.meta[bk] : 0;
            slot0 = a.meta[1 + a.nb + bk] + ((int64_t)t - tstart) * (kWave / bucket_L(WM, kid));
        }
        dispatch_bucket<M, F64, WM, 0>(a, kid, slot0, wave_id);
    }
}

// FP32 variants are held to 256 VGPRs so two waves share each SIMD (the
// guard-band redo of EDR/LCSS and ERP's gap columns would otherwise push
// them past it, to one wave per SIMD); the FP64 variants keep their full
// register file rather than spill.
#ifndef TSM_DP_WAVES_PER_EU
#define TSM_DP_WAVES_PER_EU 2
#endif
template <int M, bool F64>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(TSM_DP_WAVES_PER_EU))) void dp_kernel(DPArgs a) {
    dp_body<M, F64>(a);
}
template <int M, bool F64> __global__ __launch_bounds__(kWave) void dp_kernel_wide(DPArgs a) { dp_body<M, F64>(a); }
// STEDR / STLCSS (W <= 16 FP64 columns with timestamps) fit 256 VGPRs
// without spilling: held there, two waves share each SIMD (measured 0.56 ms
// per 100k pairs against 0.78 ms at the 272 registers the compiler picks
// unbounded, one wave per SIMD).
template <int M, bool F64>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void dp_kernel_st(DPArgs a) {
    dp_body<M, F64>(a);
}

}  // namespace tsm
