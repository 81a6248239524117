// bkma_adv_run_noalloc: bkma_adv_run for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  The global scratch that AdaptiveWg/Global and BasicRange need
// (the span3d_t global_scratch argument of bkma_run, src/core/bkma_run.hpp:10-13)
// comes from a library-owned device array instead of the caller.
//
//  - The array is `scratch` below, kBkmaStaticScratchBytes long, one copy per
//    device, held until the process ends.  Its device address is taken once
//    per device with hipGetSymbolAddress (the array's name in host code is a
//    host-side stand-in, not a device address) and kept in g_scratch_ptr.
//  - The kernels write every scratch element they read, so nothing is
//    cleared between calls.
//  - Every call on a device shares that device's array: calls that overlap on
//    one GPU (two streams, or two host threads) race.
//  - No hipMalloc, hipFree or allocating library routine is called.
#include <mutex>
#include <string>

#include "common.hpp"

static constexpr uint64_t kBkmaStaticScratchBytes = 64ull << 20;

static __device__ __align__(256) unsigned char scratch[kBkmaStaticScratchBytes];

namespace bkma {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch`.
bkma_status static_scratch_ptr(void **out) {
    int dev = 0;
    BKMA_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BKMA_ERR_UNSUPPORTED, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(BKMA_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BKMA_OK;
}

}  // namespace
}  // namespace bkma

using namespace bkma;

extern "C" uint64_t bkma_static_scratch_bytes(void) { return kBkmaStaticScratchBytes; }

extern "C" bkma_status bkma_adv_run_noalloc(bkma_impl impl, bkma_dtype dtype, void *data,
                                            const bkma_adv_params *params, const bkma_optim_params *optim,
                                            void *stream) {
    if (!data || !params) return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_adv_run_noalloc: null data or params");
    if (dtype != BKMA_F64 && dtype != BKMA_F32)
        return fail(BKMA_ERR_UNSUPPORTED, "bkma_adv_run_noalloc: dtype must be F64 or F32");
    const bool needs_scratch =
        impl == BKMA_IMPL_BASIC_RANGE || (impl == BKMA_IMPL_ADAPTIVE_WG && optim && optim->mem_space == BKMA_MEM_GLOBAL);
    bkma_span3d span{nullptr, 0, 0, 0};
    if (needs_scratch) {
        uint64_t e0 = 0, e1 = 0, e2 = 0;
        // AdvectionSolver::window() == 1 (src/solvers/AdvectionSolver.hpp:19)
        BKMA_TRY(bkma_scratch_extents(impl, optim, params->n0, params->n1, params->n2, 1, &e0, &e1, &e2));
        const uint64_t elem = dtype == BKMA_F64 ? 8 : 4;
        const unsigned __int128 bytes = (unsigned __int128)e0 * e1 * e2 * elem;
        if (bytes > kBkmaStaticScratchBytes)
            return fail(BKMA_ERR_INVALID_ARGUMENT,
                        "bkma_adv_run_noalloc: the global scratch (" + std::to_string(e0) + ", " +
                            std::to_string(e1) + ", " + std::to_string(e2) + ") of " +
                            std::to_string((unsigned long long)bytes) + " bytes exceeds the static scratch of " +
                            std::to_string(kBkmaStaticScratchBytes) + " bytes");
        BKMA_TRY(require_device());
        void *p = nullptr;
        BKMA_TRY(static_scratch_ptr(&p));
        span = bkma_span3d{p, e0, e1, e2};
    }
    return bkma_adv_run(impl, dtype, data, params, optim, span, stream);
}
