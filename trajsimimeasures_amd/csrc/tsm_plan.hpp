// tsm_plan.hpp -- the task planner of the DP kernel: a counting sort of
// pairs by (W bucket, rows), packed 64/L slots to a task.
//
//   key(pair) = (WMAX - W) * kRowKeys + (kRowKeys - 1 - min(nrow, kRowKeys - 1)),
//   W = clamp(ceil(ncol / L), 1, WMAX), nrow/ncol = longer/shorter length,
// so ascending keys give descending W, then descending rows.  Each W bucket
// is padded to whole tasks, so every task has one W; the persistent kernel
// takes tasks in key order, largest work first.
#pragma once

#include "tsm_dp.hpp"

namespace tsm {

static __global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

struct PlanArgs {
    PairSpace ps;
    const int64_t *offs;
    int64_t npairs;
    uint32_t *hist;    // [WM * kRowKeys] counts, then exclusive cursors
    uint32_t *slots;
    int64_t *wend;     // [WM]
    int64_t *ntasks;
    unsigned int *counter;
};

template <int L>
__device__ __forceinline__ int pair_bin(const PlanArgs &a, int64_t p) {
    constexpr int WM = WMax<L>::value;
    int64_t ia, ib;
    decode_pair(a.ps, p, ia, ib);
    const int64_t na = a.offs[ia + 1] - a.offs[ia];
    const int64_t nb = a.offs[ib + 1] - a.offs[ib];
    const int64_t nrow = na > nb ? na : nb;
    const int64_t ncol = na > nb ? nb : na;
    int64_t w = (ncol + L - 1) / L;
    w = w < 1 ? 1 : (w > WM ? WM : w);
    const int64_t rk = nrow < kRowKeys - 1 ? nrow : kRowKeys - 1;
    return (int)((WM - w) * kRowKeys + (kRowKeys - 1 - rk));
}

template <int L>
__global__ void plan_hist_kernel(PlanArgs a) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < a.npairs; p += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&a.hist[pair_bin<L>(a, p)], 1u);
}

// One block of 1024 threads: per W bucket, exclusive scan of its kRowKeys
// row bins; buckets are padded to whole tasks of 64/L slots.
template <int L>
__global__ __launch_bounds__(1024) void plan_scan_kernel(PlanArgs a) {
    constexpr int WM = WMax<L>::value;
    constexpr int S = kWave / L;
    static_assert(kRowKeys == 1024, "one bin per thread");
    __shared__ uint32_t sh[1024];
    __shared__ int64_t task_base;
    const int t = threadIdx.x;
    if (t == 0) task_base = 0;
    __syncthreads();
    for (int k = 0; k < WM; ++k) {
        uint32_t *h = a.hist + (int64_t)k * kRowKeys;
        const uint32_t v = h[t];
        sh[t] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
            uint32_t x = t >= o ? sh[t - o] : 0u;
            __syncthreads();
            sh[t] += x;
            __syncthreads();
        }
        const uint32_t incl = sh[t];
        const uint32_t total = sh[1023];
        const int64_t base = task_base;
        h[t] = (uint32_t)(base * S) + (incl - v);
        __syncthreads();
        if (t == 0) {
            task_base = base + (total + S - 1) / S;
            a.wend[k] = task_base;
        }
        __syncthreads();
    }
    if (t == 0) {
        *a.ntasks = task_base;
        *a.counter = 0u;
    }
}

template <int L>
__global__ void plan_scatter_kernel(PlanArgs a) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < a.npairs; p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t pos = atomicAdd(&a.hist[pair_bin<L>(a, p)], 1u);
        a.slots[pos] = (uint32_t)p;
    }
}

}  // namespace tsm
