// tsm_plan.hpp -- the task planner of the DP kernel: a counting sort of
// pairs by (bucket, rows), packed 64/L slots to a task.
//
//   bin(pair) = bucket * kRowKeys + (kRowKeys - 1 - min(nrow, 511) / 4),
//   bucket = table[ncol] (L, W) for ncol <= table range, else the strip bucket,
// with nrow/ncol the longer/shorter length.  Buckets are numbered by W
// descending, so ascending bins give descending W, then descending rows.
// Each bucket is padded to whole tasks of 64/L slots so every task has one
// (L, W); the persistent kernel takes tasks in bin order, largest work first.
//
// Three kernels: a histogram (LDS-aggregated per block), a one-block scan
// that lays the buckets out and re-zeroes the histogram for the next call,
// and a scatter that reserves each block's range per bin with one global
// atomic and ranks pairs inside it through LDS.
#pragma once

#include "tsm_dp.hpp"

namespace tsm {

constexpr int kScanThreads = 1024;
constexpr int kMaxBuckets = 64;
constexpr int kMaxBins = kMaxBuckets * kRowKeys;                   // 8192
constexpr int kBinsPerThread = kMaxBins / kScanThreads;           // 8
constexpr int kPlanBlock = 256;
constexpr int kPairsPerBlock = 1024;
static_assert(kRowKeys % kBinsPerThread == 0, "a thread's bins stay in one bucket");

struct PlanArgs {
    PairSpace ps;
    const int64_t *offs;
    int64_t npairs;
    uint32_t *hist;     // [kMaxBins] counts; zero between calls
    uint32_t *cursor;   // [kMaxBins] first slot of each bin, advanced by the scatter
    uint32_t *slots;
    int64_t *meta;      // see DPArgs::meta
    unsigned int *counter;
    int nb;
    BucketTable tab;
};

__device__ __forceinline__ int pair_bin(const PlanArgs &a, int64_t p) {
    int64_t ia, ib;
    decode_pair(a.ps, p, ia, ib);
    const int64_t na = a.offs[ia + 1] - a.offs[ia];
    const int64_t nb = a.offs[ib + 1] - a.offs[ib];
    const int64_t nrow = na > nb ? na : nb;
    const int64_t ncol = na > nb ? nb : na;
    const int b = ncol <= a.tab.ncmax ? (int)a.tab.b[ncol] : a.tab.wide;
    const int rk = (int)((nrow < 511 ? nrow : 511) >> 2);
    return b * kRowKeys + (kRowKeys - 1 - rk);
}

static __global__ __launch_bounds__(kPlanBlock) void plan_hist_kernel(PlanArgs a) {
    __shared__ uint32_t h[kMaxBins];
    const int nbins = a.nb * kRowKeys;
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock) h[i] = 0;
    __syncthreads();
    const int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock;
    const int64_t p1 = p0 + kPairsPerBlock < a.npairs ? p0 + kPairsPerBlock : a.npairs;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += kPlanBlock) atomicAdd(&h[pair_bin(a, p)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock)
        if (h[i]) atomicAdd(&a.hist[i], h[i]);
}

// One block: thread t owns bins [8t, 8t + 8) (32 threads per bucket).  A
// block scan of the per-thread sums gives each bin's position within its
// bucket; buckets are laid out one after another, each padded to whole tasks
// of 64/L slots whose unused slots are marked empty.
static __global__ __launch_bounds__(kScanThreads) void plan_scan_kernel(PlanArgs a) {
    __shared__ uint32_t sh[kScanThreads];
    __shared__ int64_t slot_base[kMaxBuckets];
    __shared__ int64_t slot_fill[kMaxBuckets];  // first padding slot of each bucket
    __shared__ int64_t slot_end[kMaxBuckets];
    constexpr int kThreadsPerBucket = kRowKeys / kBinsPerThread;
    const int t = threadIdx.x;
    const int nbins = a.nb * kRowKeys;
    uint32_t v[kBinsPerThread];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kBinsPerThread; ++i) {
        const int bin = t * kBinsPerThread + i;
        v[i] = bin < nbins ? a.hist[bin] : 0u;
        if (bin < nbins) a.hist[bin] = 0u;  // ready for the next call
        sum += v[i];
    }
    sh[t] = sum;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {  // Hillis-Steele inclusive scan
        const uint32_t x = t >= o ? sh[t - o] : 0u;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    const int b = t / kThreadsPerBucket;
    const int first = b * kThreadsPerBucket;
    const uint32_t excl_t = sh[t] - sum;
    const uint32_t excl_first = first > 0 ? sh[first - 1] : 0u;
    if (t == 0) {
        int64_t tasks = 0, slots = 0;
        for (int k = 0; k < a.nb; ++k) {
            const int last = (k + 1) * kThreadsPerBucket - 1;
            const uint32_t start = k > 0 ? sh[k * kThreadsPerBucket - 1] : 0u;
            const int64_t cnt = (int64_t)(sh[last] - start);
            const int S = kWave / bucket_L(k);
            const int64_t nt = (cnt + S - 1) / S;
            slot_base[k] = slots;
            slot_fill[k] = slots + cnt;
            slot_end[k] = slots + nt * S;
            a.meta[1 + a.nb + k] = slots;
            tasks += nt;
            slots += nt * S;
            a.meta[1 + k] = tasks;
        }
        a.meta[0] = tasks;
        *a.counter = 0u;
    }
    __syncthreads();
    if (b < a.nb) {
        uint32_t pos = (uint32_t)slot_base[b] + (excl_t - excl_first);
#pragma unroll
        for (int i = 0; i < kBinsPerThread; ++i) {
            const int bin = t * kBinsPerThread + i;
            if (bin < nbins) a.cursor[bin] = pos;
            pos += v[i];
        }
    }
    // padding slots after each bucket's pairs: at most 64/L - 1 per bucket
    if (t < a.nb * 16) {
        const int k = t / 16;
        const int64_t s = slot_fill[k] + t % 16;
        if (s < slot_end[k]) a.slots[s] = kEmptySlot;
    }
}

static __global__ __launch_bounds__(kPlanBlock) void plan_scatter_kernel(PlanArgs a) {
    __shared__ uint32_t cnt[kMaxBins];
    __shared__ uint32_t base[kMaxBins];
    const int nbins = a.nb * kRowKeys;
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock) cnt[i] = 0;
    __syncthreads();
    const int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock;
    const int64_t p1 = p0 + kPairsPerBlock < a.npairs ? p0 + kPairsPerBlock : a.npairs;
    constexpr int kPer = kPairsPerBlock / kPlanBlock;  // 4 pairs per thread
    int bins[kPer];
    uint32_t rank[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
        bins[k] = p < p1 ? pair_bin(a, p) : -1;
        rank[k] = bins[k] >= 0 ? atomicAdd(&cnt[bins[k]], 1u) : 0u;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += kPlanBlock)
        base[i] = cnt[i] ? atomicAdd(&a.cursor[i], cnt[i]) : 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t p = p0 + threadIdx.x + (int64_t)k * kPlanBlock;
        if (bins[k] >= 0) a.slots[base[bins[k]] + rank[k]] = (uint32_t)p;
    }
}

}  // namespace tsm
