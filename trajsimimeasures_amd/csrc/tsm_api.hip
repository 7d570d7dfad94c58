// tsm_api.hip -- the C ABI of libtsm.so (include/tsm.h): validation,
// device context, ERP prologue, Hausdorff launch and host-memory staging.
//
// Host flow of one batched call (the body of the reference's per-measure
// functions, e.g. core/dtw.py:56-77):
//   1. validate, pick the lanes-per-pair L (the caller's hint or a model);
//   2. plan: counting sort of pairs into tasks (tsm_plan.hpp);
//   3. one persistent DP kernel (tsm_dp.hpp; ERP's gap distances are
//      computed inside it), or the Hausdorff kernel.
// Everything is stream-ordered on the caller's stream; the only host sync
// is a workspace grow.
#include <cmath>
#include <cstring>

#include "tsm_cdds.hpp"
#include "tsm_sar.hpp"
#include "tsm_hausdorff.hpp"
#include "tsm_host.hpp"

namespace tsm {

// Interleaved (x, y) pairs: even elements subtract ox, odd ones oy.
__global__ void xy_to_f32_kernel(const double *in, float *out, int64_t n, double ox, double oy) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)(in[i] - ((i & 1) ? oy : ox));
}

thread_local std::string g_err;
std::mutex g_mu;
Context g_ctx;
// Held for the whole body of tsm_pairs_host / tsm_pairs_padded_host: their
// staging buffers (Context::h_*) are shared, so two host calls must not
// overlap.  Taken before g_mu (which tsm_pairs takes inside), never after.
std::mutex g_host_mu;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace tsm

namespace {

using namespace tsm;

// Lanes-per-pair choice: the caller's hint, else a cost model over the
// expected column count.  Waste per pair is (L - 1) skew steps over nrow rows
// plus column padding up to a multiple of L; both shrink with smaller L,
// while register columns (and so occupancy) grow.
int choose_lanes(const tsm_params &prm) {
    return prm.lanes == 2 || prm.lanes == 4 || prm.lanes == 8 || prm.lanes == 16 ? prm.lanes : 0;
}

}  // namespace

namespace tsm {

// Argument checks shared by every entry point; fills prm with defaults
// overridden by prm_in.
int validate(int measure, const tsm_trajset *set, int64_t npairs, const tsm_params *prm_in, const double *out,
             tsm_params &prm) {
    if (measure < TSM_DTW || measure > TSM_SAR) return fail(TSM_E_ARG, "unknown measure id " + std::to_string(measure));
    if (!set) return fail(TSM_E_ARG, "null trajectory set");
    if (npairs < 0) return fail(TSM_E_ARG, "negative pair count");
    if (npairs > (int64_t)0xFFFFFFFEll) return fail(TSM_E_ARG, "at most 2^32-2 pairs per call");
    if (npairs == 0) return TSM_OK;
    if (!out) return fail(TSM_E_ARG, "null output buffer");
    if (!set->offsets || set->n < 1) return fail(TSM_E_ARG, "empty trajectory set");
    if (set->max_len < 1) return fail(TSM_E_ARG, "max_len must be >= 1");
    if (set->npts < set->n) return fail(TSM_E_ARG, "npts must be >= n (every trajectory has a point)");
    tsm_params_default(&prm);
    if (prm_in) prm = *prm_in;
    const bool f64 = prm.precision == TSM_PREC_FP64;
    if (prm.precision != TSM_PREC_FP32 && prm.precision != TSM_PREC_FP64)
        return fail(TSM_E_ARG, "precision must be TSM_PREC_FP32 or TSM_PREC_FP64");
    if (prm.eps_mode != TSM_EPS_SQRT && prm.eps_mode != TSM_EPS_SQUARED)
        return fail(TSM_E_ARG, "eps_mode must be TSM_EPS_SQRT or TSM_EPS_SQUARED");
    const bool st = measure == TSM_STEDR || measure == TSM_STLCSS || measure == TSM_CDDS || measure == TSM_SAR;
    if (st && !set->t64) return fail(TSM_E_ARG, "STEDR/STLCSS/CDDS/SAR need per-point timestamps (t64)");
    if (measure == TSM_SAR) {
        if (prm.time_step < 1) return fail(TSM_E_ARG, "SAR time_step must be a positive integer");
        if (prm.target_length < 1) return fail(TSM_E_ARG, "SAR target_length must be >= 1");
        if (prm.replicas < 1) return fail(TSM_E_ARG, "SAR replicas (threads_per_traj) must be >= 1");
    }
    const bool needs64 = f64 || st || measure == TSM_EDR || measure == TSM_LCSS || measure == TSM_ERP;
    if (needs64 && !set->xy64) return fail(TSM_E_ARG, "this measure needs the FP64 coordinates (xy64)");
    if (!f64 && !st && measure != TSM_EDR && measure != TSM_LCSS && !set->xy32)
        return fail(TSM_E_ARG, "FP32 measures need the FP32 coordinates (xy32)");
    if (set->max_len > (1 << 24)) return fail(TSM_E_ARG, "trajectories longer than 2^24 points are not supported");
    return TSM_OK;
}

// The launch sequence of one call on workspace w (g_mu held).
int launch(Workspace &w, int measure, const tsm_trajset *set, const PairSpace &ps, int64_t npairs,
           const tsm_params &prm, double *out, int32_t *counts, hipStream_t st) {
    const bool f64 = prm.precision == TSM_PREC_FP64;
    if (measure == TSM_HAUSDORFF) {
        HArgs h;
        h.xy32 = set->xy32;
        h.xy64 = set->xy64;
        h.offs = set->offsets;
        h.ps = ps;
        h.npairs = npairs;
        h.nset = set->n;
        h.out = out;
        int grid = (int)(npairs < 1048576 ? npairs : 1048576);
        const ProfEvents pe;
        if (f64) hipExtLaunchKernelGGL(hausdorff_kernel<double>, dim3(grid), dim3(kWave), 0, st, pe.start, pe.stop, 0, h);
        else hipExtLaunchKernelGGL(hausdorff_kernel<float>, dim3(grid), dim3(kWave), 0, st, pe.start, pe.stop, 0, h);
        TSM_HIP(hipGetLastError());
        return TSM_OK;
    }
    if (measure == TSM_CDDS) {
        CArgs c;
        c.xy = set->xy64;
        c.t = set->t64;
        c.offs = set->offsets;
        c.ps = ps;
        c.npairs = npairs;
        c.nset = set->n;
        c.eps = prm.eps;
        c.out = out;
        int grid = (int)(npairs < 1048576 ? npairs : 1048576);
        const ProfEvents pe;
        hipExtLaunchKernelGGL(cdds_kernel, dim3(grid), dim3(kWave), 0, st, pe.start, pe.stop, 0, c);
        TSM_HIP(hipGetLastError());
        return TSM_OK;
    }
    if (measure == TSM_SAR) {
        SArgs s;
        s.xy = set->xy64;
        s.t = set->t64;
        s.offs = set->offsets;
        s.ps = ps;
        s.npairs = npairs;
        s.nset = set->n;
        s.deps = prm.eps;
        s.ts = prm.time_step;
        s.M = prm.target_length;
        s.R = prm.replicas;
        s.out = out;
        int grid = (int)(npairs < 1048576 ? npairs : 1048576);
        const ProfEvents pe;
        hipExtLaunchKernelGGL(sar_kernel, dim3(grid), dim3(kWave), 0, st, pe.start, pe.stop, 0, s);
        TSM_HIP(hipGetLastError());
        return TSM_OK;
    }
    const double thr = eps_threshold(prm.eps, prm.eps_mode);
    const int L = choose_lanes(prm);
    switch (measure) {
    case TSM_DTW:
        return f64 ? run_dp<kDTW, true>(w, L, *set, ps, npairs, thr, out, counts, st)
                   : run_dp<kDTW, false>(w, L, *set, ps, npairs, thr, out, counts, st);
    case TSM_FRECHET:
        return f64 ? run_dp<kFrechet, true>(w, L, *set, ps, npairs, thr, out, counts, st)
                   : run_dp<kFrechet, false>(w, L, *set, ps, npairs, thr, out, counts, st);
    case TSM_ERP:
        return f64 ? run_dp<kERP, true>(w, L, *set, ps, npairs, thr, out, counts, st, 0.0, prm.gx, prm.gy)
                   : run_dp<kERP, false>(w, L, *set, ps, npairs, thr, out, counts, st, 0.0, prm.gx, prm.gy);
    case TSM_EDR:
    case TSM_LCSS: {
        // FP32 match test with a guard band when the set's coordinate bound
        // gives a useful one, else the FP64 test (both exact)
        float lo, hi;
        const bool band = !f64 && set->xy32 && count_band(thr, set->coord_max, lo, hi);
        if (measure == TSM_EDR)
            return band ? run_dp<kEDR, false>(w, L, *set, ps, npairs, thr, out, counts, st)
                        : run_dp<kEDR, true>(w, L, *set, ps, npairs, thr, out, counts, st);
        return band ? run_dp<kLCSS, false>(w, L, *set, ps, npairs, thr, out, counts, st)
                    : run_dp<kLCSS, true>(w, L, *set, ps, npairs, thr, out, counts, st);
    }
    case TSM_STEDR: return run_dp<kSTEDR, true>(w, L, *set, ps, npairs, thr, out, counts, st, prm.delta);
    case TSM_STLCSS: return run_dp<kSTLCSS, true>(w, L, *set, ps, npairs, thr, out, counts, st, prm.delta);
    }
    return fail(TSM_E_ARG, "unknown measure");
}

__global__ void clear_words_kernel(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

int clear_words(uint32_t *p, int64_t n, hipStream_t st) {
    if (n <= 0) return TSM_OK;
    hipLaunchKernelGGL(clear_words_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n);
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

}  // namespace tsm

namespace {

int run(int measure, const tsm_trajset *set, const PairSpace &ps, int64_t npairs, const tsm_params *prm_in,
        double *out, int32_t *counts, void *stream) {
    tsm_params prm;
    int rc = validate(measure, set, npairs, prm_in, out, prm);
    if (rc || npairs == 0) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_ctx())) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // every call shares one workspace: a call on another stream than the
    // previous one first waits for that call's kernels
    if (!g_ctx.ws_event) TSM_HIP(hipEventCreateWithFlags(&g_ctx.ws_event, hipEventDisableTiming));
    if (g_ctx.ws_used && g_ctx.ws_stream != st) TSM_HIP(hipStreamWaitEvent(st, g_ctx.ws_event, 0));
    rc = launch(g_ctx.ws, measure, set, ps, npairs, prm, out, counts, st);
    if (rc) return rc;
    TSM_HIP(hipEventRecord(g_ctx.ws_event, st));
    g_ctx.ws_stream = st;
    g_ctx.ws_used = true;
    return TSM_OK;
}

}  // namespace

// ---------------------------------------------------------------- C ABI

extern "C" {

int tsm_abi_version(void) { return TSM_ABI_VERSION; }

const char *tsm_last_error(void) { return g_err.c_str(); }

void tsm_params_default(tsm_params *p) {
    if (!p) return;
    p->eps = 0.0;
    p->eps_mode = TSM_EPS_SQRT;
    p->gx = 0.0;
    p->gy = 0.0;
    p->lanes = 0;
    p->precision = TSM_PREC_FP32;
    p->delta = 0.0;
    p->time_step = 10;     // config.py:56 trajsimi_sar_time_eps
    p->target_length = 10; // config.py:57 trajsimi_sar_target_length
    p->replicas = 64;      // core/sar.py:248 threads_per_traj default
}

int tsm_init(int device) {
    int n = 0;
    TSM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(TSM_E_ARG, "device " + std::to_string(device) + " not present");
    TSM_HIP(hipSetDevice(device));
    std::lock_guard<std::mutex> lk(g_mu);
    return ensure_ctx();
}

int tsm_xy_to_fp32_centred(const double *xy64, float *xy32, int64_t npts, double ox, double oy, void *stream) {
    if (npts < 0 || (npts > 0 && (!xy64 || !xy32))) return fail(TSM_E_ARG, "bad point buffers");
    if (!finite64(ox) || !finite64(oy)) return fail(TSM_E_ARG, "the origin must be finite");
    if (npts == 0) return TSM_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(xy_to_f32_kernel, dim3(grid_for(2 * npts, 256)), dim3(256), 0, st, xy64, xy32, 2 * npts, ox,
                       oy);
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

int tsm_xy_to_fp32(const double *xy64, float *xy32, int64_t npts, void *stream) {
    return tsm_xy_to_fp32_centred(xy64, xy32, npts, 0.0, 0.0, stream);
}

int tsm_pairs(int measure, const tsm_trajset *set, const int32_t *pa, const int32_t *pb, int64_t npairs,
              const tsm_params *prm, double *out, int32_t *counts, void *stream) {
    if (npairs > 0 && (!pa || !pb)) return fail(TSM_E_ARG, "null pair index arrays");
    PairSpace ps{};
    ps.mode = kExplicit;
    ps.pa = pa;
    ps.pb = pb;
    return run(measure, set, ps, npairs, prm, out, counts, stream);
}

int tsm_pairs_grid(int measure, const tsm_trajset *set, int64_t q0, int64_t nq, int64_t d0, int64_t nd,
                   const tsm_params *prm, double *out, int32_t *counts, void *stream) {
    if (!set) return fail(TSM_E_ARG, "null trajectory set");
    if (nq < 0 || nd < 0 || q0 < 0 || d0 < 0 || q0 + nq > set->n || d0 + nd > set->n)
        return fail(TSM_E_ARG, "grid ranges outside the trajectory set");
    PairSpace ps{};
    ps.mode = kGrid;
    ps.q0 = q0;
    ps.nq = nq;
    ps.d0 = d0;
    ps.nd = nd;
    return run(measure, set, ps, nq * nd, prm, out, counts, stream);
}

int tsm_pairs_triu(int measure, const tsm_trajset *set, int64_t n, const tsm_params *prm, double *out,
                   int32_t *counts, void *stream) {
    if (!set) return fail(TSM_E_ARG, "null trajectory set");
    if (n < 0 || n > set->n) return fail(TSM_E_ARG, "triangle side outside the trajectory set");
    PairSpace ps{};
    ps.mode = kTriu;
    ps.n = n;
    return run(measure, set, ps, n < 2 ? 0 : n * (n - 1) / 2, prm, out, counts, stream);
}

int tsm_pairs_host(int measure, const double *xy64, const double *t64, const int64_t *offsets, int64_t n,
                   int64_t npts, const int32_t *pa, const int32_t *pb, int64_t npairs, const tsm_params *prm,
                   double *out, int32_t *counts) {
    if (!xy64 || !offsets || n < 1 || npts < 1) return fail(TSM_E_ARG, "empty trajectory set");
    if (npairs < 0) return fail(TSM_E_ARG, "negative pair count");
    if (npairs == 0) return TSM_OK;
    if (!pa || !pb || !out) return fail(TSM_E_ARG, "null pair or output array");
    if (offsets[0] != 0 || offsets[n] != npts) return fail(TSM_E_ARG, "offsets must run from 0 to npts");
    int64_t max_len = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t l = offsets[i + 1] - offsets[i];
        if (l < 1) return fail(TSM_E_ARG, "trajectory " + std::to_string(i) + " is empty");
        if (l > max_len) max_len = l;
    }
    for (int64_t k = 0; k < npairs; ++k)
        if (pa[k] < 0 || pa[k] >= n || pb[k] < 0 || pb[k] >= n)
            return fail(TSM_E_ARG, "pair " + std::to_string(k) + " indexes outside the set");
    // the FP32 copy relative to the centre of the set's bounding box; the
    // kernels are built on the premise that every input is finite
    double ox = 0.0, oy = 0.0, cmax = 0.0;
    if (fp32_origin(xy64, npts, ox, oy, cmax)) return fail(TSM_E_ARG, "coordinates must be finite");
    if (t64) {
        const uint64_t *tb = reinterpret_cast<const uint64_t *>(t64);
        for (int64_t i = 0; i < npts; ++i)
            if (!finite_bits(tb[i])) return fail(TSM_E_ARG, "timestamps must be finite");
    }
    std::lock_guard<std::mutex> host_lk(g_host_mu);
    Context &c = g_ctx;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        int rc = ensure_ctx();
        if (rc) return rc;
        if ((rc = c.h_xy64.ensure(npts * 16)) || (rc = c.h_xy32.ensure(npts * 8)) || (rc = c.h_offs.ensure((n + 1) * 8)) ||
            (rc = c.h_pa.ensure(npairs * 4)) || (rc = c.h_pb.ensure(npairs * 4)) || (rc = c.h_out.ensure(npairs * 8)) ||
            (rc = c.h_cnt.ensure(npairs * 4)) || (t64 && (rc = c.h_t64.ensure(npts * 8))))
            return rc;
    }
    TSM_HIP(hipMemcpy(c.h_xy64.p, xy64, npts * 16, hipMemcpyHostToDevice));
    if (t64) TSM_HIP(hipMemcpy(c.h_t64.p, t64, npts * 8, hipMemcpyHostToDevice));
    TSM_HIP(hipMemcpy(c.h_offs.p, offsets, (n + 1) * 8, hipMemcpyHostToDevice));
    TSM_HIP(hipMemcpy(c.h_pa.p, pa, npairs * 4, hipMemcpyHostToDevice));
    TSM_HIP(hipMemcpy(c.h_pb.p, pb, npairs * 4, hipMemcpyHostToDevice));
    int rc = tsm_xy_to_fp32_centred(c.h_xy64.as<double>(), c.h_xy32.as<float>(), npts, ox, oy, nullptr);
    if (rc) return rc;
    tsm_trajset set{c.h_xy32.as<float>(), c.h_xy64.as<double>(), c.h_offs.as<int64_t>(), n, npts, max_len, cmax,
                    t64 ? c.h_t64.as<double>() : nullptr};
    rc = tsm_pairs(measure, &set, c.h_pa.as<int32_t>(), c.h_pb.as<int32_t>(), npairs, prm, c.h_out.as<double>(),
                   counts ? c.h_cnt.as<int32_t>() : nullptr, nullptr);
    if (rc) return rc;
    TSM_HIP(hipMemcpy(out, c.h_out.p, npairs * 8, hipMemcpyDeviceToHost));
    if (counts) TSM_HIP(hipMemcpy(counts, c.h_cnt.p, npairs * 4, hipMemcpyDeviceToHost));
    return TSM_OK;
}

// Compaction of the reference's padded batches on the device: trajectory
// k < B is row k of t1 [B, L1, dim], trajectory B + k row k of t2 [B, L2,
// dim], staged one after the other in `pad`.  One wavefront per trajectory
// (grid stride); point i goes to offs[k] + i of xy64 and, when t64 is not
// null, its column 2 to t64.  Each block writes the bounding box of the
// points it copied and whether any value it read was not finite:
// box[5 b + 0..4] = lo x, lo y, hi x, hi y, 1.0 if not finite else 0.0.
constexpr int kUnpadBlock = 256;
__global__ __launch_bounds__(kUnpadBlock) void unpad_kernel(const double *pad, int64_t B, int64_t L1, int64_t L2,
                                                             int dim, const int64_t *offs, double *xy64, double *t64,
                                                             double *box) {
    constexpr int kWaves = kUnpadBlock / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * kWaves;
    for (int64_t k = (int64_t)blockIdx.x * kWaves + wave; k < 2 * B; k += stride) {
        const int64_t o = offs[k], n = offs[k + 1] - o;
        // moved as raw 64-bit words: the finite test needs the bits as loaded
        const uint64_t *src = reinterpret_cast<const uint64_t *>(pad) + (k < B ? k * L1 : B * L1 + (k - B) * L2) * dim;
        for (int64_t i = lane; i < n; i += 64) {
            const uint64_t bx = src[i * dim], by = src[i * dim + 1];
            reinterpret_cast<uint64_t *>(xy64)[2 * (o + i)] = bx;
            reinterpret_cast<uint64_t *>(xy64)[2 * (o + i) + 1] = by;
            bad |= !finite_bits(bx) || !finite_bits(by);
            const double x = __builtin_bit_cast(double, bx), y = __builtin_bit_cast(double, by);
            lx = fmin(lx, x);
            ly = fmin(ly, y);
            hx = fmax(hx, x);
            hy = fmax(hy, y);
            if (t64) {
                const uint64_t bt = src[i * dim + 2];
                reinterpret_cast<uint64_t *>(t64)[o + i] = bt;
                bad |= !finite_bits(bt);
            }
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        lx = fmin(lx, __shfl_xor(lx, m));
        ly = fmin(ly, __shfl_xor(ly, m));
        hx = fmax(hx, __shfl_xor(hx, m));
        hy = fmax(hy, __shfl_xor(hy, m));
    }
    const bool wbad = __any(bad);
    __shared__ double part[kWaves][5];
    if (lane == 0) {
        part[wave][0] = lx;
        part[wave][1] = ly;
        part[wave][2] = hx;
        part[wave][3] = hy;
        part[wave][4] = wbad ? 1.0 : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int e = threadIdx.x;
        double v = part[0][e];
        for (int w = 1; w < kWaves; ++w) v = e < 2 ? fmin(v, part[w][e]) : fmax(v, part[w][e]);
        box[5 * (int64_t)blockIdx.x + e] = v;
    }
}

__global__ void iota_pairs_kernel(int32_t *pa, int32_t *pb, int64_t B) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < B; k += (int64_t)gridDim.x * blockDim.x) {
        pa[k] = (int32_t)k;
        pb[k] = (int32_t)(B + k);
    }
}

int tsm_pairs_padded_host(int measure, const double *t1, const int64_t *len1, int64_t L1, const double *t2,
                          const int64_t *len2, int64_t L2, int64_t B, int dim, const tsm_params *prm, double *out,
                          int32_t *counts) {
    if (measure < TSM_DTW || measure > TSM_SAR) return fail(TSM_E_ARG, "unknown measure id " + std::to_string(measure));
    if (B < 0) return fail(TSM_E_ARG, "negative batch size");
    if (B == 0) return TSM_OK;
    if (B > 0x3FFFFFFFll) return fail(TSM_E_ARG, "at most 2^30 - 1 pairs per call");
    if (!t1 || !t2 || !len1 || !len2 || !out) return fail(TSM_E_ARG, "null batch, length or output array");
    if (dim != 2 && dim != 3) return fail(TSM_E_ARG, "points must have 2 (x, y) or 3 (x, y, t) values");
    const bool timed = measure == TSM_STEDR || measure == TSM_STLCSS || measure == TSM_CDDS || measure == TSM_SAR;
    if (timed && dim != 3) return fail(TSM_E_ARG, "STEDR/STLCSS/CDDS/SAR need [batch, seq, 3] points (x, y, t)");
    if (L1 < 1 || L2 < 1) return fail(TSM_E_ARG, "padded lengths must be >= 1");
    std::vector<int64_t> offs(2 * B + 1);
    offs[0] = 0;
    int64_t max_len = 0;
    for (int64_t k = 0; k < 2 * B; ++k) {
        const int64_t l = k < B ? len1[k] : len2[k - B], L = k < B ? L1 : L2;
        if (l < 1 || l > L) return fail(TSM_E_ARG, "every length must be in [1, padded length]");
        offs[k + 1] = offs[k] + l;
        if (l > max_len) max_len = l;
    }
    const int64_t npts = offs[2 * B];
    const int64_t n1 = B * L1 * dim, n2 = B * L2 * dim;
    const int nblk = (int)std::min<int64_t>((2 * B + 3) / 4, 1024);
    std::lock_guard<std::mutex> host_lk(g_host_mu);
    Context &c = g_ctx;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        int rc = ensure_ctx();
        if (rc) return rc;
        if ((rc = c.h_pad.ensure((n1 + n2) * 8)) || (rc = c.h_box.ensure(nblk * 40)) ||
            (rc = c.h_xy64.ensure(npts * 16)) || (rc = c.h_xy32.ensure(npts * 8)) ||
            (rc = c.h_offs.ensure((2 * B + 1) * 8)) || (rc = c.h_pa.ensure(B * 4)) || (rc = c.h_pb.ensure(B * 4)) ||
            (rc = c.h_out.ensure(B * 8)) || (rc = c.h_cnt.ensure(B * 4)) || (timed && (rc = c.h_t64.ensure(npts * 8))))
            return rc;
    }
    double *pad = c.h_pad.as<double>();
    TSM_HIP(hipMemcpy(pad, t1, n1 * 8, hipMemcpyHostToDevice));
    TSM_HIP(hipMemcpy(pad + n1, t2, n2 * 8, hipMemcpyHostToDevice));
    TSM_HIP(hipMemcpy(c.h_offs.p, offs.data(), (2 * B + 1) * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(unpad_kernel, dim3(nblk), dim3(kUnpadBlock), 0, nullptr, pad, B, L1, L2, dim,
                       c.h_offs.as<int64_t>(), c.h_xy64.as<double>(), timed ? c.h_t64.as<double>() : nullptr,
                       c.h_box.as<double>());
    TSM_HIP(hipGetLastError());
    hipLaunchKernelGGL(iota_pairs_kernel, dim3(grid_for(B, 256)), dim3(256), 0, nullptr, c.h_pa.as<int32_t>(),
                       c.h_pb.as<int32_t>(), B);
    TSM_HIP(hipGetLastError());
    std::vector<double> box(5 * (size_t)nblk);
    TSM_HIP(hipMemcpy(box.data(), c.h_box.p, box.size() * 8, hipMemcpyDeviceToHost));
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int b = 0; b < nblk; ++b) {
        if (box[5 * b + 4] != 0.0) return fail(TSM_E_ARG, "coordinates and timestamps must be finite");
        lo[0] = std::fmin(lo[0], box[5 * b]);
        lo[1] = std::fmin(lo[1], box[5 * b + 1]);
        hi[0] = std::fmax(hi[0], box[5 * b + 2]);
        hi[1] = std::fmax(hi[1], box[5 * b + 3]);
    }
    double ox = 0.0, oy = 0.0, cmax = 0.0;
    fp32_origin_of_box(lo, hi, ox, oy, cmax);
    int rc = tsm_xy_to_fp32_centred(c.h_xy64.as<double>(), c.h_xy32.as<float>(), npts, ox, oy, nullptr);
    if (rc) return rc;
    tsm_trajset set{c.h_xy32.as<float>(), c.h_xy64.as<double>(), c.h_offs.as<int64_t>(), 2 * B, npts, max_len, cmax,
                    timed ? c.h_t64.as<double>() : nullptr};
    rc = tsm_pairs(measure, &set, c.h_pa.as<int32_t>(), c.h_pb.as<int32_t>(), B, prm, c.h_out.as<double>(),
                   counts ? c.h_cnt.as<int32_t>() : nullptr, nullptr);
    if (rc) return rc;
    TSM_HIP(hipMemcpy(out, c.h_out.p, B * 8, hipMemcpyDeviceToHost));
    if (counts) TSM_HIP(hipMemcpy(counts, c.h_cnt.p, B * 4, hipMemcpyDeviceToHost));
    return TSM_OK;
}

int tsm_tune_set_lanes(int coord_bits, const uint8_t *lanes_of_ncol, int n) {
    if (coord_bits != 32 && coord_bits != 64) return fail(TSM_E_ARG, "coord_bits must be 32 or 64");
    const int k = coord_bits == 64 ? 1 : 0;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!lanes_of_ncol) {
        g_ctx.tuned[k] = false;
        return TSM_OK;
    }
    if (n < 2 || n > 257) return fail(TSM_E_ARG, "lanes table must cover column counts 1..n-1 with 2 <= n <= 257");
    for (int i = 1; i < n; ++i)
        if (lanes_of_ncol[i] != 2 && lanes_of_ncol[i] != 4 && lanes_of_ncol[i] != 8 && lanes_of_ncol[i] != 16)
            return fail(TSM_E_ARG, "lanes must be 2, 4, 8 or 16");
    std::memset(g_ctx.tuned_lanes[k], 0, sizeof(g_ctx.tuned_lanes[k]));
    std::memcpy(g_ctx.tuned_lanes[k], lanes_of_ncol, (size_t)n);
    g_ctx.tuned[k] = true;
    return TSM_OK;
}

int tsm_tune_set_waves_per_cu(int waves_per_cu) {
    if (waves_per_cu < 0 || waves_per_cu > 32) return fail(TSM_E_ARG, "waves_per_cu must be in 0..32 (0 = default 8)");
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx.waves_per_cu = waves_per_cu == 0 ? 8 : waves_per_cu;
    return TSM_OK;
}

int tsm_tune_set_wide_lanes(int lanes) {
    if (lanes != 0 && lanes != 16 && lanes != 64) return fail(TSM_E_ARG, "wide lanes must be 0 (by launch size), 16 or 64");
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx.wide_lanes = lanes;
    return TSM_OK;
}

int tsm_tune_set_direct_pairs(int64_t max_pairs) {
    if (max_pairs < -1) return fail(TSM_E_ARG, "direct pairs must be -1 (default), 0 (never) or a pair count");
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx.direct_pairs = max_pairs;
    return TSM_OK;
}

}  // extern "C"

// A captured batch: the planner and kernel launches of one call, replayable
// on the same device buffers, with a private workspace.
struct tsm_graph {
    tsm::Workspace ws;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Capture of one call over pair space ps: a first eager run sizes the
// graph's private workspace (fresh buffers are zeroed there, outside the
// capture) and produces the results once; the same launches are then
// recorded.  The planner leaves its histogram and counters zero behind
// itself, so replays need no clears and the graph holds no memset nodes.
int capture(int measure, const tsm_trajset *set, const PairSpace &ps, int64_t npairs, const tsm_params *prm_in,
            double *out, int32_t *counts, void *stream, tsm_graph **graph) {
    if (!graph) return fail(TSM_E_ARG, "null graph handle");
    *graph = nullptr;
    if (!stream) return fail(TSM_E_ARG, "graph capture needs a non-null stream");
    if (npairs < 1) return fail(TSM_E_ARG, "graph capture needs at least one pair");
    tsm_params prm;
    int rc = validate(measure, set, npairs, prm_in, out, prm);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_ctx())) return rc;
    tsm_graph *g = new tsm_graph();
    if ((rc = launch(g->ws, measure, set, ps, npairs, prm, out, counts, st))) {
        g->ws.release();
        delete g;
        return rc;
    }
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        g_ctx.capturing = true;
        rc = launch(g->ws, measure, set, ps, npairs, prm, out, counts, st);
        g_ctx.capturing = false;
        hipError_t e2 = hipStreamEndCapture(st, &g->graph);
        if (rc == TSM_OK && e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess && rc == TSM_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess || rc != TSM_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->ws.release();
        delete g;
        return rc != TSM_OK ? rc : fail(TSM_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    *graph = g;
    return TSM_OK;
}

}  // namespace

extern "C" {

int tsm_graph_capture_pairs(int measure, const tsm_trajset *set, const int32_t *pa, const int32_t *pb,
                            int64_t npairs, const tsm_params *prm_in, double *out, int32_t *counts, void *stream,
                            tsm_graph **graph) {
    if (npairs > 0 && (!pa || !pb)) return fail(TSM_E_ARG, "null pair index arrays");
    PairSpace ps{};
    ps.mode = kExplicit;
    ps.pa = pa;
    ps.pb = pb;
    return capture(measure, set, ps, npairs, prm_in, out, counts, stream, graph);
}

int tsm_graph_capture_grid(int measure, const tsm_trajset *set, int64_t q0, int64_t nq, int64_t d0, int64_t nd,
                           const tsm_params *prm_in, double *out, int32_t *counts, void *stream, tsm_graph **graph) {
    if (!set) return fail(TSM_E_ARG, "null trajectory set");
    if (nq < 0 || nd < 0 || q0 < 0 || d0 < 0 || q0 + nq > set->n || d0 + nd > set->n)
        return fail(TSM_E_ARG, "grid ranges outside the trajectory set");
    PairSpace ps{};
    ps.mode = kGrid;
    ps.q0 = q0;
    ps.nq = nq;
    ps.d0 = d0;
    ps.nd = nd;
    return capture(measure, set, ps, nq * nd, prm_in, out, counts, stream, graph);
}

int tsm_graph_capture_triu(int measure, const tsm_trajset *set, int64_t n, const tsm_params *prm_in, double *out,
                           int32_t *counts, void *stream, tsm_graph **graph) {
    if (!set) return fail(TSM_E_ARG, "null trajectory set");
    if (n < 0 || n > set->n) return fail(TSM_E_ARG, "triangle side outside the trajectory set");
    PairSpace ps{};
    ps.mode = kTriu;
    ps.n = n;
    return capture(measure, set, ps, n < 2 ? 0 : n * (n - 1) / 2, prm_in, out, counts, stream, graph);
}

int tsm_graph_launch(tsm_graph *graph, void *stream) {
    if (!graph || !graph->exec) return fail(TSM_E_ARG, "null graph");
    TSM_HIP(hipGraphLaunch(graph->exec, reinterpret_cast<hipStream_t>(stream)));
    return TSM_OK;
}

int tsm_graph_destroy(tsm_graph *graph) {
    if (!graph) return TSM_OK;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    graph->ws.release();
    delete graph;
    return TSM_OK;
}

int tsm_profile_enable(int every) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (every < 0) return fail(TSM_E_ARG, "profile stride must be >= 0");
    g_ctx.prof = every != 0;
    g_ctx.prof_every = every > 0 ? every : 1;
    g_ctx.prof_seen = 0;
    g_ctx.ev_used = 0;
    return TSM_OK;
}

int tsm_profile_read(double *total_ms, int64_t *launches) {
    std::lock_guard<std::mutex> lk(g_mu);
    double ms = 0.0;
    for (size_t i = 0; i < g_ctx.ev_used; ++i) {
        TSM_HIP(hipEventSynchronize(g_ctx.ev[i].second));
        float t = 0.f;
        TSM_HIP(hipEventElapsedTime(&t, g_ctx.ev[i].first, g_ctx.ev[i].second));
        ms += t;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = (int64_t)g_ctx.ev_used;
    g_ctx.ev_used = 0;
    return TSM_OK;
}

}  // extern "C"
