// tsm_noalloc.hip -- tsm_pairs_noalloc: tsm_pairs without any device
// allocation, for callers that cannot allocate (code under stream capture,
// serving loops that forbid hipMalloc).
//
// The planner's histogram, fill counters, task slots and queue counters and
// the column-strip boundary vectors of one call are carved from `scratch`, a
// fixed device array of kNoAllocBytes (a direct DP launch and the Hausdorff,
// CDDS and SAR kernels take none of it).
// Each device holds its own copy of the array for the life of the process,
// at its own address: the host finds it with hipGetSymbolAddress on the
// entry point's first call on a device (with that device current) and uses
// only that address afterwards.  Every call on a device shares the array and
// finds it as the previous call left it, so the planner's zero-start
// counters are cleared by a fill kernel at the start of each call, and two
// calls that overlap on one device (different streams, or threads without
// ordering) race on it.
#include "tsm_host.hpp"

namespace {

using namespace tsm;

constexpr size_t kNoAllocBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` per device ordinal (nullptr until looked up).
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (g_mu held).
int scratch_of_current_device(void **ptr) {
    int dev = 0;
    TSM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TSM_E_UNSUP, "device ordinal beyond the no-allocation table");
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(TSM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return TSM_OK;
}

}  // namespace

extern "C" int tsm_pairs_noalloc(int measure, const tsm_trajset *set, const int32_t *pa, const int32_t *pb,
                                 int64_t npairs, const tsm_params *prm_in, double *out, int32_t *counts,
                                 void *stream) {
    if (npairs > 0 && (!pa || !pb)) return fail(TSM_E_ARG, "null pair index arrays");
    tsm_params prm;
    int rc = validate(measure, set, npairs, prm_in, out, prm);
    if (rc || npairs == 0) return rc;
    PairSpace ps{};
    ps.mode = kExplicit;
    ps.pa = pa;
    ps.pb = pb;
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_ctx())) return rc;
    void *base = nullptr;
    if ((rc = scratch_of_current_device(&base))) return rc;
    Arena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = kNoAllocBytes;
    Workspace w;
    w.arena = &arena;
    const bool was = g_ctx.capturing;
    g_ctx.capturing = true;  // no profiling events (hipEventCreate) either
    rc = launch(w, measure, set, ps, npairs, prm, out, counts, reinterpret_cast<hipStream_t>(stream));
    g_ctx.capturing = was;
    return rc;
}
