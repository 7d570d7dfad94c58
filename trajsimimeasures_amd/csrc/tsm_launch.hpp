// tsm_launch.hpp -- host launch sequence of the DP path: plan kernels, then
// the persistent DP kernel.  Included by the per-measure translation units,
// which instantiate run_dp for their measure.
#pragma once

#include "tsm_host.hpp"
#include "tsm_plan.hpp"

namespace tsm {

template <int M, bool F64, int L>
int run_dp_L(const tsm_trajset &set, const PairSpace &ps, int64_t npairs, double thr, double *out,
             int32_t *counts, hipStream_t st) {
    constexpr int WM = WMax<L>::value;
    constexpr int S = kWave / L;
    using AT = typename Rec<M, F64>::AT;
    Context &c = g_ctx;
    const int64_t nbins = (int64_t)WM * kRowKeys;
    const int64_t nslots = npairs + (int64_t)WM * S;
    int rc;
    if ((rc = c.hist.ensure(nbins * 4)) || (rc = c.slots.ensure(nslots * 4)) || (rc = c.wend.ensure(WM * 8)) ||
        (rc = c.ntasks.ensure(8)) || (rc = c.counter.ensure(4)))
        return rc;

    auto kern = dp_kernel<M, F64, L>;
    int grid = persistent_grid(kern);
    const int64_t max_tasks = (npairs + S - 1) / S + WM;
    if (grid > max_tasks) grid = (int)max_tasks;
    int64_t bnd_stride = 0;
    if (set.max_len > (int64_t)L * WM) {  // some pair may need column strips
        bnd_stride = set.max_len;
        if ((rc = c.bnd.ensure((size_t)grid * S * bnd_stride * sizeof(AT)))) return rc;
    }

    PlanArgs pa;
    pa.ps = ps;
    pa.offs = set.offsets;
    pa.npairs = npairs;
    pa.hist = c.hist.as<uint32_t>();
    pa.slots = c.slots.as<uint32_t>();
    pa.wend = c.wend.as<int64_t>();
    pa.ntasks = c.ntasks.as<int64_t>();
    pa.counter = c.counter.as<unsigned int>();

    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(nbins, 256)), dim3(256), 0, st, pa.hist, nbins, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(nslots, 256)), dim3(256), 0, st, pa.slots, nslots, kEmptySlot);
    hipLaunchKernelGGL(plan_hist_kernel<L>, dim3(grid_for(npairs, 256)), dim3(256), 0, st, pa);
    hipLaunchKernelGGL(plan_scan_kernel<L>, dim3(1), dim3(1024), 0, st, pa);
    hipLaunchKernelGGL(plan_scatter_kernel<L>, dim3(grid_for(npairs, 256)), dim3(256), 0, st, pa);
    TSM_HIP(hipGetLastError());

    DPArgs a;
    a.xy32 = set.xy32;
    a.xy64 = set.xy64;
    a.offs = set.offsets;
    a.gap32 = c.gap32.as<float>();
    a.gap64 = c.gap64.as<double>();
    a.gsum32 = c.gsum32.as<float>();
    a.gsum64 = c.gsum64.as<double>();
    a.ps = ps;
    a.slots = pa.slots;
    a.wend = pa.wend;
    a.ntasks = pa.ntasks;
    a.counter = pa.counter;
    a.bnd = c.bnd.p;
    a.bnd_stride = bnd_stride;
    a.thr = thr;
    a.out = out;
    a.counts = counts;
    {
        ProfScope ps_(st);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kWave), 0, st, a);
    }
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

template <int M, bool F64>
int run_dp(int L, const tsm_trajset &set, const PairSpace &ps, int64_t npairs, double thr, double *out,
           int32_t *counts, hipStream_t st) {
    switch (L) {
    case 8: return run_dp_L<M, F64, 8>(set, ps, npairs, thr, out, counts, st);
    default: return run_dp_L<M, F64, 16>(set, ps, npairs, thr, out, counts, st);
    }
}

}  // namespace tsm
