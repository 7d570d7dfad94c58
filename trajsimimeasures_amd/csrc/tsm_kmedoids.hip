// tsm_kmedoids.hip -- FasterPAM k-medoids over a device dissimilarity
// matrix, the clustering step of test_clustering_time.py:353-367
// (kmedoids.fasterpam; E. Schubert, P. J. Rousseeuw, "Fast and eager
// k-medoids clustering: O(k) runtime improvement of the PAM, CLARA, and
// CLARANS algorithms", Information Systems 101, 2021), and the expansion of
// the all-pairs upper triangle into the symmetric matrix it clusters
// (test_clustering_time.py:256-269).
//
// State per point (structure of arrays): nearest medoid (position in the
// medoid list, distance) and second nearest.  FasterPAM scans candidates in
// index order and swaps eagerly, so candidate j's evaluation depends on every
// swap before it.  The GPU evaluates a batch of consecutive candidates at
// once against the current state (one wavefront per candidate: lanes stride
// over the points reading the candidate's matrix row, with per-lane partial
// sums of the shared gain and of the per-medoid removal-term changes), then
// takes the first candidate whose best swap lowers the loss.  Candidates up
// to it saw exactly the state the sequential algorithm gives them; after the
// swap the scan resumes behind it.  The batch doubles while no swap is found
// and shrinks after one.
//
// Sums over the points use fixed orders.  A candidate's: partial l adds the
// points l, l + 64, ... in ascending order, then partials combine pairwise
// (l + 32 into l, ..., l + 1 into l: a butterfly whose lane 0 holds this
// tree).  The loss and removal losses: per block of 256 points, then blocks in
// order (km_block_partials).  The matrix is HBM-bound input: each candidate
// streams its row once.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tsm.h"
#include "../../include/tsm_cluster.h"
#include "tsm_host.hpp"

namespace tsm {
namespace {

constexpr int kKmMax = 64;
constexpr uint32_t kKmNone = 0xFFFFFFFFu;

struct KmState {
    uint32_t *ni, *si;  // nearest / second nearest medoid (position in the medoid list)
    double *nd, *sd;    // their distances
};

// lane 0 gets the fixed-order tree of the 64 lanes' partials
__device__ __forceinline__ double km_tree(double v) {
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Loss and removal losses (sum of nearest distances; per medoid, sum of
// second - nearest over its points) in their fixed order: the point kernels
// run blocks of 256 points, one point per thread; each block reduces its
// k + 1 values as the trees of its four 64-point groups (wave butterflies)
// combined as (g0 + g2) + (g1 + g3) into part[block][v] (v = k: the loss);
// km_blocksum_kernel then adds the blocks in order.  Terms of other medoids
// are +0, which never changes a sum here (no partial is ever -0).
constexpr int kKmBlock = 256;

__device__ __forceinline__ void km_block_partials(int k, bool ok, uint32_t ni, double nd, double sd, double *part) {
    __shared__ double g[4][kKmMax + 1];
    const int t = threadIdx.x, w = t >> 6, l = t & (kWave - 1);
    for (int v = 0; v <= k; ++v) {
        double x = 0.0;
        if (ok) x = v == k ? nd : (ni == (uint32_t)v ? sd - nd : 0.0);
        x = km_tree(x);
        if (l == 0) g[w][v] = x;
    }
    __syncthreads();
    for (int v = t; v <= k; v += kKmBlock) part[(int64_t)blockIdx.x * (k + 1) + v] = (g[0][v] + g[2][v]) + (g[1][v] + g[3][v]);
}

// One wavefront: rl[m] = sum over blocks of part[b][m] in block order, loss likewise.
// With ctl: only after a pending swap, which it then clears.
__global__ __launch_bounds__(kWave) void km_blocksum_kernel(const double *part, int64_t nblk, int k, double *rl,
                                                            double *loss, int64_t *pend) {
    if (pend && *pend < 0) return;
    for (int v = threadIdx.x; v <= k; v += kWave) {
        double s = 0.0;
        constexpr int U = 16;
        for (int64_t b0 = 0; b0 < nblk; b0 += U) {
            double x[U];
#pragma unroll
            for (int u = 0; u < U; ++u) x[u] = part[(b0 + u < nblk ? b0 + u : nblk - 1) * (k + 1) + v];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (b0 + u < nblk) s += x[u];
        }
        if (v < k) rl[v] = s;
        else *loss = s;
    }
    if (pend && threadIdx.x == 0) *pend = -1;
}

__global__ __launch_bounds__(kKmBlock) void km_assign_kernel(const double *D, int64_t n, int k, const int64_t *med,
                                                              KmState s, double *part) {
    const int64_t i = blockIdx.x * (int64_t)kKmBlock + threadIdx.x;
    const bool ok = i < n;
    uint32_t ni = 0, si = kKmNone;
    double nd = 0.0, sd = 0.0;
    if (ok) {
        nd = D[i * n + med[0]];
        for (int m = 1; m < k; ++m) {
            const double d = D[i * n + med[m]];
            if (d < nd || i == med[m]) {  // a medoid is its own nearest
                si = ni;
                sd = nd;
                ni = (uint32_t)m;
                nd = d;
            } else if (si == kKmNone || d < sd) {
                si = (uint32_t)m;
                sd = d;
            }
        }
        s.ni[i] = ni;
        s.si[i] = si;
        s.nd[i] = nd;
        s.sd[i] = sd;
    }
    km_block_partials(k, ok, ni, nd, sd, part);
}

// The scan's control block, kept on the device so that the host enqueues
// whole rounds (evaluate a batch, pick, swap, sums) without waiting on each
// batch's outcome.  Every kernel of a round reads it; only km_pick_kernel
// (one wavefront, between the batch's evaluation and the swap) writes it,
// and km_blocksum_kernel clears the pending swap.
struct KmCtl {
    int64_t j;         // first candidate of the next batch
    int64_t end;       // the pass scans candidates [0, end): the last swap's candidate, n after a swap
    int64_t batch;     // current batch size
    int64_t cnt;       // candidates in the next batch
    int64_t lastswap;  // candidate of the last swap (n before any)
    int64_t swaps;
    int64_t iter;      // passes started
    int64_t before;    // swaps when the current pass started
    int64_t max_iter;
    int64_t pend_b;    // pending swap: medoid position (-1: none) ...
    int64_t pend_c;    // ... and the candidate that replaces it
    int64_t done;
};
constexpr int64_t kKmMinBatch = 512;

// Candidates ctl.j + i, i < ctl.cnt (blocks loop over i): loss change of the
// best swap bringing candidate j in and the medoid it replaces (+inf for a
// current medoid).
__global__ __launch_bounds__(kWave) void km_eval_kernel(const double *D, int64_t n, int k, KmState s,
                                                        const double *rl, const int64_t *med, const KmCtl *ctl,
                                                        double *change, int32_t *best) {
    extern __shared__ double q[];  // k * 64 doubles, lane-private columns
    if (ctl->done) return;
    const int l = threadIdx.x;
    const int64_t j0 = ctl->j, cnt = ctl->cnt;
    for (int64_t i = blockIdx.x; i < cnt; i += gridDim.x) {
        const int64_t j = j0 + i;
        if (j == med[s.ni[j]]) {
            if (l == 0) {
                change[i] = __builtin_huge_val();
                best[i] = 0;
            }
            continue;
        }
        for (int m = 0; m < k; ++m) q[m * kWave + l] = 0.0;
        const double *row = D + j * n;
        double acc = 0.0;
        constexpr int U = 16;  // points per lane whose loads are in flight together
        for (int64_t base = l; base < n; base += U * kWave) {
            double dj[U], nd[U], sd[U];
            uint32_t ni[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {  // unconditional loads (index clamped in bounds; unused past n)
                const int64_t o = base + u * kWave < n ? base + u * kWave : n - 1;
                dj[u] = row[o];
                nd[u] = s.nd[o];
                sd[u] = s.sd[o];
                ni[u] = s.ni[o];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // j would be o's nearest: shared gain, and n(o)'s removal term is void; j would be o's
                // second: n(o)'s removal term shrinks.  Branch-free: otherwise (and past n) both terms
                // are +0, which leaves the partials as they are (they are never -0: every term is <= 0).
                const bool ok = base + u * kWave < n;
                const bool nearer = ok && dj[u] < nd[u], second = ok && !nearer && dj[u] < sd[u];
                acc += nearer ? dj[u] - nd[u] : 0.0;
                q[ni[u] * kWave + l] += nearer ? nd[u] - sd[u] : (second ? dj[u] - sd[u] : 0.0);
            }
        }
        acc = km_tree(acc);
        double bl = 0.0;
        int b = 0;
        for (int m = 0; m < k; ++m) {
            const double v = rl[m] + km_tree(q[m * kWave + l]);
            if (m == 0 || v < bl) {
                bl = v;
                b = m;
            }
        }
        if (l == 0) {
            change[i] = bl + acc;
            best[i] = b;
        }
    }
}

// The first candidate of the batch whose change is negative becomes the
// pending swap (its medoid position updated in med); then the scan moves on
// exactly as the sequential algorithm's: behind the swap with a halved batch,
// or past the batch with a doubled one; a pass that reaches its end starts
// the next (ending at the last swap) unless it made no swap or max_iter
// passes ran.
__global__ __launch_bounds__(kWave) void km_pick_kernel(const double *change, const int32_t *best, int64_t n,
                                                        int64_t *med, KmCtl *ctl) {
    if (ctl->done) return;
    const int l = threadIdx.x;
    const int64_t j0 = ctl->j, cnt = ctl->cnt;
    int64_t f = -1;
    for (int64_t c0 = 0; c0 < cnt; c0 += kWave) {
        const int64_t c = c0 + l;
        const bool neg = c < cnt && change[c] < 0.0;
        const uint64_t mask = __builtin_amdgcn_ballot_w64(neg);
        if (mask) {
            f = c0 + __builtin_ctzll(mask);
            break;
        }
    }
    if (l != 0) return;
    const int64_t minb = kKmMinBatch < n ? kKmMinBatch : n;
    KmCtl c = *ctl;
    if (f >= 0) {
        const int64_t cand = j0 + f;
        const int b = best[f];
        med[b] = cand;
        c.pend_b = b;
        c.pend_c = cand;
        ++c.swaps;
        c.lastswap = cand;
        c.end = n;
        c.j = cand + 1;
        c.batch = c.batch / 2 > minb ? c.batch / 2 : minb;
    } else {
        c.j += cnt;
        c.batch = 2 * c.batch < n ? 2 * c.batch : n;
    }
    while (!c.done && c.j >= c.end) {  // the pass is over
        if (c.swaps == c.before || c.iter >= c.max_iter) {
            c.done = 1;
        } else {
            ++c.iter;
            c.before = c.swaps;
            c.end = c.lastswap;
            c.j = 0;
        }
    }
    c.cnt = c.done ? 0 : (c.batch < c.end - c.j ? c.batch : c.end - c.j);
    *ctl = c;
}

__device__ __forceinline__ void km_second(const double *D, int64_t n, int k, const int64_t *med, uint32_t nn,
                                          uint32_t b, int64_t o, double djo, uint32_t &si, double &sd) {
    uint32_t s = b;
    double d = djo;
    for (int i = 0; i < k; ++i) {
        if ((uint32_t)i == nn || (uint32_t)i == b) continue;
        const double e = D[o * n + med[i]];
        if (e < d) {
            s = (uint32_t)i;
            d = e;
        }
    }
    si = s;
    sd = d;
}

// Medoid b is now point j (med already updated): every point's nearest and
// second nearest.
__global__ __launch_bounds__(kKmBlock) void km_swap_kernel(const double *D, int64_t n, int k, const int64_t *med,
                                                            KmState s, const KmCtl *ctl, double *part) {
    if (ctl->pend_b < 0) return;
    const uint32_t b = (uint32_t)ctl->pend_b;
    const int64_t j = ctl->pend_c;
    const int64_t o = blockIdx.x * (int64_t)kKmBlock + threadIdx.x;
    const bool ok = o < n;
    uint32_t ni = 0, si = 0;
    double nd = 0.0, sd = 0.0;
    if (ok) {
        ni = s.ni[o];
        si = s.si[o];
        nd = s.nd[o];
        sd = s.sd[o];
        if (o == j) {
            if (ni != b) {
                si = ni;
                sd = nd;
            }
            ni = b;
            nd = 0.0;
        } else {
            const double djo = D[j * n + o];
            if (ni == b) {  // its nearest medoid left
                if (djo < sd) {
                    nd = djo;
                } else {
                    ni = si;
                    nd = sd;
                    km_second(D, n, k, med, ni, b, o, djo, si, sd);
                }
            } else if (djo < nd) {
                si = ni;
                sd = nd;
                ni = b;
                nd = djo;
            } else if (djo < sd) {
                si = b;
                sd = djo;
            } else if (si == b) {  // its second nearest left
                km_second(D, n, k, med, ni, b, o, djo, si, sd);
            }
        }
        s.ni[o] = ni;
        s.si[o] = si;
        s.nd[o] = nd;
        s.sd[o] = sd;
    }
    km_block_partials(k, ok, ni, nd, sd, part);
}

// Row sums without the diagonal, fixed order (k = 1).
__global__ __launch_bounds__(kWave) void km_rowsum_kernel(const double *D, int64_t n, double *sums) {
    const int l = threadIdx.x;
    const int64_t j = blockIdx.x;
    double p = 0.0;
    for (int64_t o = l; o < n; o += kWave)
        if (o != j) p += D[j * n + o];
    p = km_tree(p);
    if (l == 0) sums[j] = p;
}

__global__ void km_labels_kernel(const uint32_t *ni, int64_t n, int32_t *labels) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        labels[i] = (int32_t)ni[i];
}

__global__ void square_from_triu_kernel(const double *tri, int64_t n, double *D) {
    const int64_t total = n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / n, j = e - i * n;
        double v = 0.0;
        if (i < j) v = tri[triu_row_start(i, n) + (j - i - 1)];
        else if (j < i) v = tri[triu_row_start(j, n) + (i - j - 1)];
        D[e] = v;
    }
}

struct KmBuffers {
    DevBuf ni, si, nd, sd, rl, loss, med, change, best, ctl, sums, part;
    int64_t *h_pin = nullptr;  // pinned: [0, 16) the control block, [16] loss, [32, 32 + 64) medoids
};
constexpr int kKmPinWords = 96;
static_assert(sizeof(KmCtl) <= 16 * sizeof(int64_t), "control block in the pinned words [0, 16)");

// Rounds enqueued between two reads of the control block: a round after the
// scan ended costs four no-op launches, a read one host round trip.
constexpr int kKmRoundsPerCheck = 16;
std::mutex g_km_mu;
KmBuffers g_km;

int km_grid(int64_t n) {
    const int64_t g = (n + 255) / 256;
    return (int)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

}  // namespace
}  // namespace tsm

using namespace tsm;

extern "C" int tsm_square_from_triu(const double *tri, int64_t n, double *diss, void *stream) {
    if (n < 1) return fail(TSM_E_ARG, "square_from_triu: n must be >= 1");
    if (!diss || (n > 1 && !tri)) return fail(TSM_E_ARG, "square_from_triu: null buffer");
    if (n > (int64_t)3037000499ll) return fail(TSM_E_ARG, "square_from_triu: n too large");
    const hipStream_t st = (hipStream_t)stream;
    const int64_t total = n * n;
    const int64_t g = (total + 255) / 256;
    hipLaunchKernelGGL(square_from_triu_kernel, dim3((unsigned)(g < 1048576 ? g : 1048576)), dim3(256), 0, st, tri, n,
                       diss);
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

extern "C" int tsm_kmedoids_fasterpam(const double *diss, int64_t n, int32_t k, int64_t *medoids, int32_t max_iter,
                                      int32_t *labels, double *loss, int32_t *n_iter, int32_t *n_swap, void *stream) {
    if (!diss || !medoids) return fail(TSM_E_ARG, "kmedoids: null matrix or medoid array");
    if (n < 1 || n > 0x7FFFFFFFll) return fail(TSM_E_ARG, "kmedoids: n must be in [1, 2^31)");
    if (k < 1 || k > kKmMax || k > n) return fail(TSM_E_ARG, "kmedoids: k must be in [1, min(n, 64)]");
    if (max_iter < 0) return fail(TSM_E_ARG, "kmedoids: max_iter must be >= 0");
    {
        std::vector<int64_t> seen(medoids, medoids + k);
        for (int i = 0; i < k; ++i) {
            if (medoids[i] < 0 || medoids[i] >= n) return fail(TSM_E_ARG, "kmedoids: initial medoid outside [0, n)");
            for (int j = 0; j < i; ++j)
                if (seen[j] == medoids[i]) return fail(TSM_E_ARG, "kmedoids: initial medoids must be distinct");
        }
    }
    const hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_km_mu);
    KmBuffers &w = g_km;
    int rc;
    if ((rc = w.med.ensure(sizeof(int64_t) * k)) || (rc = w.ctl.ensure(sizeof(KmCtl)))) return rc;
    if (!w.h_pin) TSM_HIP(hipHostMalloc((void **)&w.h_pin, kKmPinWords * sizeof(int64_t), hipHostMallocDefault));
    KmCtl *h_ctl = reinterpret_cast<KmCtl *>(w.h_pin);
    int64_t *h_med = w.h_pin + 32;
    double *h_lossp = reinterpret_cast<double *>(w.h_pin + 16);
    std::vector<int64_t> med(medoids, medoids + k);
    double h_loss = 0.0;
    int iter = 0, swaps = 0;

    if (k == 1) {  // the point with the least distance sum (the initial one unless another is strictly less)
        if ((rc = w.sums.ensure(sizeof(double) * n))) return rc;
        hipLaunchKernelGGL(km_rowsum_kernel, dim3((unsigned)n), dim3(kWave), 0, st, diss, n, w.sums.as<double>());
        TSM_HIP(hipGetLastError());
        std::vector<double> sums(n);
        TSM_HIP(hipStreamSynchronize(st));
        TSM_HIP(hipMemcpy(sums.data(), w.sums.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        const int64_t first = med[0];
        int64_t bestj = first;
        double sb = sums[first];
        for (int64_t j = 0; j < n; ++j) {
            if (j == first) continue;
            if (sums[j] < sb) {
                bestj = j;
                sb = sums[j];
            }
        }
        med[0] = bestj;
        h_loss = sb;
        iter = 1;
        swaps = bestj != first;
        if (labels) TSM_HIP(hipMemsetAsync(labels, 0, sizeof(int32_t) * n, st));
    } else {
        if ((rc = w.ni.ensure(4 * n)) || (rc = w.si.ensure(4 * n)) || (rc = w.nd.ensure(8 * n)) ||
            (rc = w.sd.ensure(8 * n)) || (rc = w.rl.ensure(8 * k)) || (rc = w.loss.ensure(8)) ||
            (rc = w.change.ensure(8 * n)) || (rc = w.best.ensure(4 * n)))
            return rc;
        const int64_t nblk = (n + kKmBlock - 1) / kKmBlock;
        if ((rc = w.part.ensure(sizeof(double) * nblk * (k + 1)))) return rc;
        double *part = w.part.as<double>();
        KmState s{w.ni.as<uint32_t>(), w.si.as<uint32_t>(), w.nd.as<double>(), w.sd.as<double>()};
        int64_t *dmed = w.med.as<int64_t>();
        const size_t lds = sizeof(double) * kWave * (size_t)k;
        std::memcpy(h_med, med.data(), sizeof(int64_t) * k);
        TSM_HIP(hipMemcpyAsync(dmed, h_med, sizeof(int64_t) * k, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(km_assign_kernel, dim3((unsigned)nblk), dim3(kKmBlock), 0, st, diss, n, (int)k, dmed, s,
                           part);
        hipLaunchKernelGGL(km_blocksum_kernel, dim3(1), dim3(kWave), 0, st, part, nblk, (int)k, w.rl.as<double>(),
                           w.loss.as<double>(), nullptr);
        TSM_HIP(hipGetLastError());
        // The scan runs on the device (km_pick_kernel steers it through the
        // control block): the host enqueues rounds of evaluate / pick / swap /
        // sums and reads the control block once per kKmRoundsPerCheck rounds
        // instead of each batch's outcome.
        KmCtl c0{};
        c0.j = 0;
        c0.end = n;
        c0.batch = kKmMinBatch < n ? kKmMinBatch : n;
        c0.cnt = c0.batch;
        c0.lastswap = n;
        c0.max_iter = max_iter;
        c0.iter = max_iter > 0 ? 1 : 0;
        c0.done = max_iter > 0 ? 0 : 1;
        c0.pend_b = -1;
        *h_ctl = c0;
        KmCtl *dctl = w.ctl.as<KmCtl>();
        TSM_HIP(hipMemcpyAsync(dctl, h_ctl, sizeof(KmCtl), hipMemcpyHostToDevice, st));
        const int eval_grid = (int)(n < 4096 ? n : 4096);  // blocks loop over a batch's candidates
        // every round moves the scan past at least one candidate, so a pass
        // takes at most n rounds: a scan that has not ended by then is a fault
        const double round_cap = ((double)max_iter + 1.0) * ((double)n + 1.0) + kKmRoundsPerCheck;
        double rounds = 0.0;
        while (true) {
            if (rounds > round_cap) return fail(TSM_E_HIP, "kmedoids: the device scan did not finish");
            rounds += kKmRoundsPerCheck;
            for (int r = 0; r < kKmRoundsPerCheck; ++r) {
                hipLaunchKernelGGL(km_eval_kernel, dim3((unsigned)eval_grid), dim3(kWave), lds, st, diss, n, (int)k, s,
                                   w.rl.as<double>(), dmed, dctl, w.change.as<double>(), w.best.as<int32_t>());
                hipLaunchKernelGGL(km_pick_kernel, dim3(1), dim3(kWave), 0, st, w.change.as<double>(),
                                   w.best.as<int32_t>(), n, dmed, dctl);
                hipLaunchKernelGGL(km_swap_kernel, dim3((unsigned)nblk), dim3(kKmBlock), 0, st, diss, n, (int)k, dmed,
                                   s, dctl, part);
                hipLaunchKernelGGL(km_blocksum_kernel, dim3(1), dim3(kWave), 0, st, part, nblk, (int)k,
                                   w.rl.as<double>(), w.loss.as<double>(), &dctl->pend_b);
            }
            TSM_HIP(hipGetLastError());
            TSM_HIP(hipMemcpyAsync(h_ctl, dctl, sizeof(KmCtl), hipMemcpyDeviceToHost, st));
            TSM_HIP(hipStreamSynchronize(st));
            if (h_ctl->done) break;
        }
        iter = (int)h_ctl->iter;
        swaps = (int)h_ctl->swaps;
        TSM_HIP(hipMemcpyAsync(h_med, dmed, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
        TSM_HIP(hipMemcpyAsync(h_lossp, w.loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
        if (labels)
            hipLaunchKernelGGL(km_labels_kernel, dim3(km_grid(n)), dim3(256), 0, st, w.ni.as<uint32_t>(), n, labels);
        TSM_HIP(hipGetLastError());
        TSM_HIP(hipStreamSynchronize(st));
        h_loss = *h_lossp;
        for (int i = 0; i < k; ++i) med[i] = h_med[i];
    }
    for (int i = 0; i < k; ++i) medoids[i] = med[i];
    if (loss) *loss = h_loss;
    if (n_iter) *n_iter = iter;
    if (n_swap) *n_swap = swaps;
    return TSM_OK;
}
