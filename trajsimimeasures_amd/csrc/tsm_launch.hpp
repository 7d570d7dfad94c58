// tsm_launch.hpp -- host launch sequence of the DP path: bucket table, plan
// kernels, then the persistent DP kernel.  Included by the per-measure
// translation units, which instantiate run_dp for their measure.
#pragma once

#include <cstring>
#include <utility>

#include "tsm_host.hpp"
#include "tsm_plan.hpp"

namespace tsm {

// Whether an L-lane segment holds nc columns in registers.
inline bool lanes_fit(int WM, int L, int nc) {
    return (L == 2 || L == 4 || L == 8 || L == 16) && (nc + L - 1) / L <= bucket_wmax(WM, L);
}

// Cost-model lanes for nc columns: per pair, a segment of L lanes takes
// (nrow + L - 1) steps of (c0 + c1 * W) issue slots for 64/L pairs at once,
// so the per-pair cost is L * (nrow + L - 1) * (c0 + c1 * W) / 64 with
// c0 ~ c1 (per-step bookkeeping ~ one cell) and nrow >= nc (taken as
// max(nc, 32)): the fewest lanes whose W fits, nearly always.
inline int model_lanes(int WM, int nc) {
    const double rows = nc > 32 ? nc : 32;
    int best = 16;
    double bc = 0.0;
    for (int L = 16; L >= 2; L >>= 1) {
        if (!lanes_fit(WM, L, nc)) continue;
        const double c = L * (rows + L - 1) * (1.0 + (nc + L - 1) / L);
        if (L == 16 || c <= bc) {
            best = L;
            bc = c;
        }
    }
    return best;
}

// Column count -> (L, W) variant, W = ceil(nc / L): by priority, the
// caller's lanes hint, the autotuned lanes table (tsm_tune_set_lanes), then
// the cost model, each where its L holds the columns (16 lanes always do
// within the table).  Pairs wider than the table: 16-lane segments (strips
// of 16 * WM columns), or with wide_lanes = 64, 16 lanes up to one such
// strip and a whole wavefront beyond (W = 12, 16, 20 or 24: the smallest
// that holds the pair, then W = WM in strips of 64 * WM).
// Long rows: the choice above prices a pair as if its rows were about its
// columns.  A pair with far more rows is a chain of rows + L - 1 steps of W
// cells that no other wave can shorten, and in a launch of few pairs such
// chains outlast the rest of the work (1,000 queries x 100 candidates of
// lognormal lengths: 0.7 % of the pairs had over 500 rows, and the launch
// took 0.47 ms against 0.31 ms without them).  So in the launches that use
// 64-lane segments (no lanes hint), a pair with more than kLongRows rows
// takes 16 lanes (W = ceil(nc / 16)) up to kLongRowCols16 columns and a
// whole wavefront (W = 4, 8, ...) beyond: for 900 x 30 points a chain of
// 915 steps of 2 cells instead of 901 of 15, for 900 x 300 one of 963
// steps of 8 cells instead of 915 of 19, at most doubling that pair's lane
// work.
constexpr int kLongRows = 384;

// Only the variants that the set's lengths (shorter sides up to max_len)
// can reach become buckets of the launch, numbered in schedule order: the
// 64-lane ones, then by W and L descending (per step, a task's time grows
// with W; per task, its steps with L); with long rows, the 16-lane buckets
// before the narrower ones (they hold the long chains).
inline BucketTable bucket_table(int WM, int coord_bits, int lanes_hint, int wide_lanes, int64_t max_len) {
    BucketTable t{};
    t.ncmax = 16 * WM < kTabMax - 1 ? 16 * WM : kTabMax - 1;
    t.wide16 = 16 * WM;
    t.wm = WM;
    t.wide64 = wide_lanes == 64;
    const int k = coord_bits == 64 ? 1 : 0;
    const uint8_t *tuned = g_ctx.tuned[k] ? g_ctx.tuned_lanes[k] : nullptr;
    bool used[kMaxBuckets] = {};
    uint8_t var[kTabMax] = {};
    for (int nc = 1; nc <= t.ncmax; ++nc) {
        int L;
        if (lanes_fit(WM, lanes_hint, nc)) L = lanes_hint;
        else if (tuned && lanes_fit(WM, tuned[nc], nc)) L = tuned[nc];
        else L = model_lanes(WM, nc);
        var[nc] = (uint8_t)bucket_of(WM, L, (nc + L - 1) / L);
        if (nc <= max_len || nc == 1) used[var[nc]] = true;
    }
    t.long_rows = (t.wide64 && lanes_hint == 0 && max_len > kLongRows) ? kLongRows : 0;
    if (t.long_rows) {
        for (int W = 1; W <= kLongRowCols16 / 16 && 16 * (W - 1) < max_len; ++W) used[bucket_of(WM, 16, W)] = true;
        for (int64_t nc = kLongRowCols16 + 1; nc <= max_len && nc <= t.wide16; nc = ((nc - 1) / 256 + 1) * 256 + 1)
            used[bucket_w64(WM, nc)] = true;  // each 64-lane W up to wide16 columns
    }
    const int wide_id = bucket_of(WM, 16, WM);
    if (max_len > t.ncmax) {
        used[wide_id] = true;
        if (t.wide64 && max_len > t.wide16)
            for (int64_t nc = t.wide16 + 1; nc <= max_len; nc = ((nc - 1) / 256 + 1) * 256 + 1) {  // each 64-lane W
                used[bucket_w64(WM, nc)] = true;
                if (nc > 64 * (int64_t)WM) break;
            }
    }
    int order[kMaxBuckets], n = 0;
    for (int b = 0; b < bucket_count(WM); ++b)
        if (used[b]) order[n++] = b;
    const bool l16_first = t.long_rows != 0;
    auto before = [WM, l16_first](int x, int y) {
        const int lx = bucket_L(WM, x), ly = bucket_L(WM, y);
        if ((lx == 64) != (ly == 64)) return lx == 64;
        if (l16_first && (lx == 16) != (ly == 16)) return lx == 16;
        if (bucket_W(WM, x) != bucket_W(WM, y)) return bucket_W(WM, x) > bucket_W(WM, y);
        return lx > ly;
    };
    for (int i = 1; i < n; ++i)  // insertion sort: at most kMaxBuckets entries
        for (int j = i; j > 0 && before(order[j], order[j - 1]); --j) std::swap(order[j], order[j - 1]);
    uint8_t num[kMaxBuckets] = {};
    for (int i = 0; i < n; ++i) {
        num[order[i]] = (uint8_t)i;
        t.kid[i] = (uint8_t)order[i];
        t.lanes[i] = (uint8_t)bucket_L(WM, order[i]);
    }
    t.nb = n;
    // column counts beyond max_len occur only with an understated max_len
    // (their pairs land in bucket 0 and are computed or reported unfit there)
    for (int nc = 1; nc <= t.ncmax; ++nc) t.b[nc] = used[var[nc]] ? num[var[nc]] : 0;
    t.b[0] = t.b[1];
    t.wide = used[wide_id] ? num[wide_id] : 0;
    for (int i = 0; i < bucket_n64(WM); ++i) t.w64[i] = used[i] ? num[i] : 0;
    if (t.long_rows) {
        for (int W = 1; W <= kLongRowCols16 / 16; ++W) {
            const int id = bucket_of(WM, 16, W);
            t.w16[W] = used[id] ? num[id] : 0;
        }
        t.w16[0] = t.w16[1];
    }
    return t;
}

constexpr int64_t kSmallLaunchPairs = 1 << 20;
// Direct launches by default up to three rounds of the persistent grid's
// resident waves: measured on MI355X with batches of 20..200-point pairs
// (DTW, EDR, ERP, Frechet), a direct launch takes 0.40-0.45x the planned
// one's time at 512 pairs, 0.70-0.76x at 4,096 and 1.09-1.17x at 8,192.
constexpr int64_t kDirectRounds = 3;

template <int M, bool F64>
int run_dp(Workspace &c, int lanes_hint, const tsm_trajset &set, const PairSpace &ps, int64_t npairs, double thr,
           double *out, int32_t *counts, hipStream_t st, double delta, double gx, double gy) {
    using CT = typename Rec<M, F64>::CT;
    using AT = typename Rec<M, F64>::AT;
    constexpr int WM = Rec<M, F64>::kWMax;
    static_assert(bucket_count(WM) <= kMaxBuckets && WM % 4 == 0 && WM >= 16 && WM / 4 <= 8, "bucket count");
    void (*kern)(DPArgs);
    if constexpr (Rec<M, F64>::kTime) kern = dp_kernel_st<M, F64>;
    else if constexpr (F64) kern = dp_kernel_wide<M, F64>;
    else kern = dp_kernel<M, F64>;

    DPArgs a{};
    a.xy32 = set.xy32;
    a.xy64 = set.xy64;
    a.offs = set.offsets;
    a.gx = gx;
    a.gy = gy;
    a.ps = ps;
    a.nset = set.n;
    a.thr = thr;
    a.t64 = set.t64;
    a.delta = delta;
    a.thr_lo = 0.0f;
    a.thr_hi = 0.0f;
    if constexpr (Rec<M, F64>::kBand) {
        if (!count_band(thr, set.coord_max, a.thr_lo, a.thr_hi))
            return fail(TSM_E_ARG, "internal: FP32 match test chosen without a valid guard band");
    }
    a.out = out;
    a.counts = counts;

    // Few pairs: one wavefront per pair, no planner (a launch that size
    // lasts as long as its longest pair's chain of steps, and the planner's
    // two kernels would cost about as much as the pairs themselves).
    const int64_t direct_max = g_ctx.direct_pairs >= 0
                                   ? g_ctx.direct_pairs
                                   : kDirectRounds * (int64_t)g_ctx.num_cu * g_ctx.waves_per_cu;
    if (lanes_hint == 0 && npairs <= direct_max && npairs <= 0x7fffffff && set.max_len <= 64 * (int64_t)WM) {
        a.slots = nullptr;  // direct: wave w computes pair w, its row trajectory in dynamic LDS
        const ProfEvents pe;
        hipExtLaunchKernelGGL(kern, dim3((unsigned)npairs), dim3(kWave), (uint32_t)direct_lds_bytes<M, F64>(), st, pe.start,
                              pe.stop, 0, a);
        TSM_HIP(hipGetLastError());
        return TSM_OK;
    }

    // A launch of few pairs lasts as long as its longest pair's chain of
    // steps: there, pairs wider than the table take a whole wavefront (a
    // quarter of the steps of 16 lanes, at more lane-steps per pair); in a
    // large launch the total work sets the time and 16-lane strips use the
    // lanes better.
    const int wide_lanes = g_ctx.wide_lanes ? g_ctx.wide_lanes : (npairs <= kSmallLaunchPairs ? 64 : 16);
    const BucketTable tab = bucket_table(WM, sizeof(CT) * 8, lanes_hint, wide_lanes, set.max_len);
    const int NB = tab.nb;
    const int64_t nslots = npairs + (int64_t)NB * (kWave / 2);  // each bucket padded to whole tasks
    int rc;
    // hist holds the planner's histogram and then its fill counters
    if ((rc = c.get(c.hist, 2 * kMaxBins * 4)) || (rc = c.get(c.slots, nslots * 4)) ||
        (rc = c.get(c.meta, (1 + 2 * NB) * 8)) || (rc = c.get(c.counter, 8)))
        return rc;

    int grid = persistent_grid((npairs + 3) / 4 + NB);
    int64_t bnd_stride = 0;
    if (set.max_len > tab.ncmax) {  // some pair may need column strips
        bnd_stride = set.max_len;
        const size_t per_wave = (size_t)kMaxSeg * bnd_stride * sizeof(AT);
        // a fixed arena holds the boundary columns of fewer resident waves
        // (the kernel is persistent: any grid drains the queue)
        if (c.arena && (size_t)grid * per_wave > c.arena_left()) {
            const size_t fit = c.arena_left() / per_wave;
            if (fit < 1)
                return fail(TSM_E_ARG, "trajectories of " + std::to_string(set.max_len) +
                                           " points need more boundary scratch than the no-allocation entry point holds");
            grid = (int)fit;
        }
        if ((rc = c.get(c.bnd, (size_t)grid * per_wave))) return rc;
    }

    PlanArgs pa;
    pa.ps = ps;
    pa.offs = set.offsets;
    pa.npairs = npairs;
    pa.hist = c.hist.as<uint32_t>();
    pa.fill = c.hist.as<uint32_t>() + kMaxBins;
    pa.slots = c.slots.as<uint32_t>();
    pa.meta = c.meta.as<int64_t>();
    pa.counter = c.counter.as<unsigned int>();
    pa.nset = set.n;
    pa.out = out;
    pa.counts = counts;
    pa.nb = NB;
    pa.tab = tab;

    a.slots = pa.slots;
    a.meta = pa.meta;
    a.nb = NB;
    std::memcpy(a.kid, tab.kid, sizeof(a.kid));
    a.counter = pa.counter;
    a.bnd = c.bnd.p;
    a.bnd_stride = bnd_stride;
    a.plan_zero = pa.hist;
    a.plan_zero_n = 2 * kMaxBins;

    // The planner expects zero counters: fresh buffers start zeroed and every
    // DP launch re-zeroes them behind its planner; arena memory holds
    // whatever the last call left, so it is cleared first.  Every check that
    // can fail ran above: once the planner is queued its DP kernel follows.
    if (c.arena && (rc = clear_words(pa.hist, 2 * kMaxBins, st))) return rc;
    const int pblocks = (int)((npairs + kPairsPerBlock - 1) / kPairsPerBlock);
    hipLaunchKernelGGL(plan_hist_kernel, dim3(pblocks), dim3(kPlanBlock), 0, st, pa);
    hipLaunchKernelGGL(plan_scatter_kernel, dim3(pblocks), dim3(kPlanBlock), 0, st, pa);
    const ProfEvents pe;
    hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kWave), 0, st, pe.start, pe.stop, 0, a);
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

}  // namespace tsm
