// tsm_plan.hpp -- the task planner of the DP kernel: a counting sort of
// pairs by (bucket, rows), packed 64/L slots to a task.
//
//   bin(pair) = bucket * kRowKeys + (kRowKeys - 1 - min(nrow, 511) * kRowKeys / 512),
//   bucket = table[ncol] (L, W) for ncol <= table range, else the strip bucket,
// with nrow/ncol the longer/shorter length.  Buckets are numbered by W
// descending, so ascending bins give descending W, then descending rows.
// Each bucket is padded to whole tasks of 64/L slots so every task has one
// (L, W); the persistent kernel takes tasks in bin order, largest work first.
//
// Two kernels: a histogram (LDS-aggregated per block, one global atomic per
// bin a block touches), and a scatter in which every block lays the buckets
// out itself from that histogram (a block scan in LDS: no block waits for
// another and no serial layout step sits between the kernels), reserves its
// range in each bin with one global atomic on the bin's fill counter and
// ranks its pairs inside it through LDS.  Block 0 also writes the task
// table and the padding slots.  Nothing here waits on another block: the
// histogram and the fill counters are re-zeroed for the next call by the
// first waves of the DP kernel that follows (a kernel boundary orders it
// after every read here), since a device-scope fence per block to find the
// last one costs more than the whole scatter on gfx950.
#pragma once

#include "tsm_dp.hpp"

namespace tsm {

constexpr int kMaxBins = kMaxBuckets * kRowKeys;                   // 12288
constexpr int kPlanBlock = 384;  // four threads per bucket in the layout scan
// Pairs per planner block.  Measured on 100k pairs (DTW step minus kernel):
// 256 adds ~22 us and 512 ~5 us to the planner; 2048 is the same as 1024.
#ifndef TSM_PAIRS_PER_BLOCK
#define TSM_PAIRS_PER_BLOCK 1536
#endif
constexpr int kPairsPerBlock = TSM_PAIRS_PER_BLOCK;
static_assert(kPairsPerBlock % kPlanBlock == 0, "pairs per planner block: a multiple of the block size");

struct PlanArgs {
    PairSpace ps;
    const int64_t *offs;
    int64_t npairs;
    uint32_t *hist;     // [kMaxBins] pairs per bin; zero between calls
    uint32_t *fill;     // [kMaxBins] pairs of each bin placed so far; zero between calls
    uint32_t *slots;
    int64_t *meta;      // see DPArgs::meta
    unsigned int *counter;
    int64_t nset;        // trajectories in the set: explicit pair indices must lie in [0, nset)
    double *out;         // results: NaN for a pair with an index outside the set
    int32_t *counts;     // EDR/LCSS counts (or null): -1 for such a pair
    int nb;
    BucketTable tab;
};

__device__ __forceinline__ int bin_of(const PlanArgs &a, int64_t na, int64_t nb) {
    const int64_t nrow = na > nb ? na : nb;
    const int64_t ncol = na > nb ? nb : na;
    int b;
    if (a.tab.long_rows && nrow > a.tab.long_rows && ncol <= a.tab.wide16)
        b = ncol <= kLongRowCols16 ? (int)a.tab.w16[(ncol + 15) / 16] : (int)a.tab.w64[bucket_w64(a.tab.wm, ncol)];
    else if (ncol <= a.tab.ncmax) b = (int)a.tab.b[ncol];
    else if (!a.tab.wide64 || ncol <= a.tab.wide16) b = a.tab.wide;
    else b = a.tab.w64[bucket_w64(a.tab.wm, ncol)];
    const int rk = (int)((nrow < 511 ? nrow : 511) * kRowKeys / 512);
    return b * kRowKeys + (kRowKeys - 1 - rk);
}

// Bins of this thread's pairs p0 + threadIdx.x + k * kPlanBlock (-1 past
// p1).  The pair lookups of all kPer pairs are issued before any is waited
// for, then the offset lookups, so a thread pays two memory round trips
// rather than two per pair.
constexpr int kPer = kPairsPerBlock / kPlanBlock;  // pairs per thread

// bins[k] = -1 past p1, kBadPair for an explicit pair with an index outside
// the set (it is never scheduled; the scatter writes its NaN result).
constexpr int kBadPair = -2;

__device__ __forceinline__ void pair_bins(const PlanArgs &a, int64_t p0, int64_t p1, int (&bins)[kPer]) {
    int64_t ia[kPer], ib[kPer];
    bool bad[kPer];
    if (a.ps.mode == kExplicit) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
            const int64_t pc = p < p1 ? p : p0;
            ia[k] = a.ps.pa[pc];
            ib[k] = a.ps.pb[pc];
        }
    } else {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
            decode_pair(a.ps, p < p1 ? p : p0, ia[k], ib[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        bad[k] = ia[k] < 0 || ia[k] >= a.nset || ib[k] < 0 || ib[k] >= a.nset;
        if (bad[k]) ia[k] = ib[k] = 0;
    }
    int64_t na[kPer], nb[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        na[k] = a.offs[ia[k] + 1] - a.offs[ia[k]];
        nb[k] = a.offs[ib[k] + 1] - a.offs[ib[k]];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
        bins[k] = p >= p1 ? -1 : bad[k] ? kBadPair : bin_of(a, na[k], nb[k]);
    }
}

// The bucket layout, computed by every scatter block from the global
// histogram: base[bin] = the first slot of the bin.  Thread t owns the
// kScanBins bins [kScanBins * (t % 4), + kScanBins) of bucket t / 4.  A
// block scan of the per-thread sums gives each bin's position within its
// bucket; buckets are laid out one after another, each padded to whole tasks
// of 64/L slots whose unused slots are marked empty.  The block passed
// `writer` also writes the task table (meta), the padding slots and resets
// the task counter.
constexpr int kScanTPB = kPlanBlock / kMaxBuckets;  // threads per bucket
constexpr int kScanBins = kRowKeys / kScanTPB;      // bins per thread
static_assert(kScanTPB * kMaxBuckets == kPlanBlock && kScanBins * kScanTPB == kRowKeys, "scan layout");

// The histogram words thread threadIdx.x owns in the layout, loaded ahead of
// the layout so the loads overlap the pair lookups.
__device__ __forceinline__ void layout_load(const PlanArgs &a, uint32_t (&v)[kScanBins]) {
    const int b = threadIdx.x / kScanTPB;
    const int bin0 = b * kRowKeys + (threadIdx.x % kScanTPB) * kScanBins;
#pragma unroll
    for (int i = 0; i < kScanBins; ++i) v[i] = b < a.nb ? a.hist[bin0 + i] : 0u;
}

__device__ __forceinline__ void plan_layout(const PlanArgs &a, uint32_t *base, bool writer,
                                            const uint32_t (&v)[kScanBins]) {
    __shared__ uint32_t wsum[kPlanBlock / kWave];
    __shared__ uint32_t incl[kPlanBlock];
    __shared__ int64_t slot_base[kMaxBuckets];
    __shared__ int64_t slot_fill[kMaxBuckets];  // first padding slot of each bucket
    __shared__ int64_t slot_end[kMaxBuckets];
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
    const int b = t / kScanTPB;
    const int bin0 = b * kRowKeys + (t % kScanTPB) * kScanBins;
    const bool live = b < a.nb;
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanBins; ++i) sum += v[i];
    // inclusive scan of `sum`: within each wavefront, then across the four
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[wave] = x;
    __syncthreads();
    for (int w = 0; w < wave; ++w) x += wsum[w];
    incl[t] = x;
    __syncthreads();
    const int first = b * kScanTPB;
    const uint32_t excl_t = x - sum;
    const uint32_t excl_first = first > 0 ? incl[first - 1] : 0u;
    if (wave == 0) {
        // per bucket (one lane each, 64 buckets a pass): pairs, tasks of
        // 64/L slots, then an inclusive scan of tasks and slots across the
        // buckets, carried from pass to pass
        uint32_t tcarry = 0, scarry = 0;
        for (int h = 0; h < kMaxBuckets; h += kWave) {
            const int k = h + lane;
            uint32_t cnt = 0, nt = 0, ns = 0;
            if (k < a.nb) {
                const uint32_t start = k > 0 ? incl[k * kScanTPB - 1] : 0u;
                cnt = incl[(k + 1) * kScanTPB - 1] - start;
                const uint32_t S = (uint32_t)(kWave / a.tab.lanes[k]);
                nt = (cnt + S - 1) / S;
                ns = nt * S;
            }
            uint32_t ti = nt, si = ns;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t tu = __shfl_up(ti, o, kWave), su = __shfl_up(si, o, kWave);
                if (lane >= o) {
                    ti += tu;
                    si += su;
                }
            }
            ti += tcarry;
            si += scarry;
            if (k < a.nb) {
                const int64_t sb = (int64_t)(si - ns);
                slot_base[k] = sb;
                slot_fill[k] = sb + cnt;
                slot_end[k] = (int64_t)si;
                if (writer) {
                    a.meta[1 + a.nb + k] = sb;
                    a.meta[1 + k] = (int64_t)ti;
                    if (k == a.nb - 1) a.meta[0] = (int64_t)ti;
                }
            }
            tcarry = __shfl(ti, kWave - 1, kWave);
            scarry = __shfl(si, kWave - 1, kWave);
        }
        if (writer && lane == 0) *a.counter = 0u;
    }
    __syncthreads();
    if (live) {
        uint32_t pos = (uint32_t)slot_base[b] + (excl_t - excl_first);
#pragma unroll
        for (int i = 0; i < kScanBins; ++i) {
            base[bin0 + i] = pos;
            pos += v[i];
        }
    }
    if (writer) {
        // padding slots after each bucket's pairs: at most 64/L - 1 < 32 per bucket
        for (int q = t; q < a.nb * 32; q += kPlanBlock) {
            const int k = q / 32;
            const int64_t s = slot_fill[k] + q % 32;
            if (s < slot_end[k]) a.slots[s] = kEmptySlot;
        }
    }
}

static __global__ __launch_bounds__(kPlanBlock) void plan_hist_kernel(PlanArgs a) {
    __shared__ uint32_t h[kMaxBins];
    const int nbins = a.nb * kRowKeys;
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock) h[i] = 0;
    __syncthreads();
    const int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock;
    const int64_t p1 = p0 + kPairsPerBlock < a.npairs ? p0 + kPairsPerBlock : a.npairs;
    int bins[kPer];
    pair_bins(a, p0, p1, bins);
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (bins[k] >= 0) atomicAdd(&h[bins[k]], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock)
        if (h[i]) atomicAdd(&a.hist[i], h[i]);
}

static __global__ __launch_bounds__(kPlanBlock) void plan_scatter_kernel(PlanArgs a) {
    __shared__ uint32_t cnt[kMaxBins];
    __shared__ uint32_t base[kMaxBins];
    const int nbins = a.nb * kRowKeys;
    uint32_t hv[kScanBins];
    layout_load(a, hv);  // in flight during the pair lookups
    const int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock;
    const int64_t p1 = p0 + kPairsPerBlock < a.npairs ? p0 + kPairsPerBlock : a.npairs;
    int bins[kPer];
    uint32_t rank[kPer];
    pair_bins(a, p0, p1, bins);
    plan_layout(a, base, blockIdx.x == 0, hv);
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock) cnt[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        rank[k] = bins[k] >= 0 ? atomicAdd(&cnt[bins[k]], 1u) : 0u;
        if (bins[k] == kBadPair) {
            const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
            a.out[p] = __builtin_bit_cast(double, 0x7ff8000000000000ull);  // quiet NaN
            if (a.counts) a.counts[p] = -1;
        }
    }
    __syncthreads();
    // reserve this block's range in every bin it uses: all of a thread's
    // atomics are issued before any result is waited for
    constexpr int kBinsPerThread = kMaxBins / kPlanBlock;
    uint32_t got[kBinsPerThread];
#pragma unroll
    for (int k = 0; k < kBinsPerThread; ++k) {
        const int i = threadIdx.x + k * kPlanBlock;
        const uint32_t c = i < nbins ? cnt[i] : 0u;
        got[k] = c ? atomicAdd(&a.fill[i], c) : 0u;
    }
#pragma unroll
    for (int k = 0; k < kBinsPerThread; ++k) {
        const int i = threadIdx.x + k * kPlanBlock;
        if (i < nbins) base[i] += got[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
        if (bins[k] >= 0) a.slots[base[bins[k]] + rank[k]] = (uint32_t)p;
    }
}

}  // namespace tsm
