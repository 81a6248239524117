// Header-only HIP executors of the BKMA interface, generic over the solver:
//   AdaptiveWg  src/core/AdaptiveWg.hpp:57-140  (LDS or caller-global scratch)
//   NDRange     src/core/NDRange.hpp:4-40       (one work-group per line)
//   BasicRange  src/core/BasicRange.hpp:27-69   (out of place, then a copy)
//   bkma_run    src/core/bkma_run.hpp:10-62     (host batch loop)
// A solver is any by-value device functor with
//   i64 window() const                                (host and device)
//   double operator()(const Line &line, i64 i0, i64 i1, i64 i2) const  (device)
// where line(i) is the value of row i of the (i0, :, i2) line and the result
// is the new value of row i1 (stored at i1 - (window() - 1)), the contract of
// the reference's solvers (AdvectionSolver.hpp:71-98, ConvSolver.hpp:20-44).
// The library builds its AdvectionSolver and ConvSolver kernels from these
// templates; a caller's own solver, compiled with hipcc for gfx950, runs
// through them with bkma::bkma_run (bkma/plugin.hpp).
//
// Three defects of the reference's AdaptiveWg kernel are not reproduced
// (src/core/AdaptiveWg.hpp:108-134):
//  - a work-item's line loop ends at the end of its batch, so no line of the
//    next batch is stepped twice when the batch size is not a multiple of the
//    grid stride;
//  - every work-item of a work-group runs the same number of line iterations
//    (out-of-range lines skip the loads and stores but keep the barriers), so
//    no barrier is ever divergent;
//  - a second barrier after the copy-back keeps a work-group from computing
//    its next line into the scratch while peers still copy the previous one.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "bkma_amd.h"

namespace bkma {
namespace exec {

using i64 = long long;
using u64 = unsigned long long;

// Outcome of a host call: a status code and, on failure, the message.
struct Status {
    bkma_status code = BKMA_OK;
    std::string msg;
    bool ok() const { return code == BKMA_OK; }
};
inline Status error(bkma_status c, std::string m) { return Status{c, std::move(m)}; }
inline Status hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return Status{};
    return error(BKMA_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
inline Status launch_status(const char *what) { return hip_status(hipGetLastError(), what); }
#define BKMA_EXEC_TRY(expr)              \
    do {                                 \
        ::bkma::exec::Status _s = (expr); \
        if (!_s.ok()) return _s;         \
    } while (0)

inline u64 div_up(u64 a, u64 b) { return (a + b - 1) / b; }

// One strided (i0, :, i2) line of the array as the solver sees it.
template <class T> struct Line {
    T *base;
    i64 stride;
    __device__ __forceinline__ double operator()(i64 i) const { return (double)base[i * stride]; }
    __device__ __forceinline__ T &at(i64 i) const { return base[i * stride]; }
};

// ---------------------------------------------------------------------------
// AdaptiveWg.  Work-group (w2, w1, w0) on HIP (x, y, z): x is the
// fastest-varying id, as dim 2 is in the reference's SYCL range.  The
// reference's 3D grid (g0*w0, w1, g2*w2) is flattened onto
// blockIdx.x = bg2 + g2 * bg0.
// ---------------------------------------------------------------------------
struct AwgLaunch {
    i64 n0, n1, n2;
    i64 b0_size, b0_offset, b2_size, b2_offset;
    i64 w0, w1, w2, g0, g2;
    i64 trips0, trips2;  // uniform per-work-group line iterations
    i64 ge1, ge2;        // global scratch extents 1 and 2 (Global only)
    i64 i0_base;         // slab index of the launch's first slab (data and the solver's i0)
};

template <class T, class Solver, bool GLOBAL_SCRATCH>
__global__ void adaptivewg_kernel(T *__restrict__ data, Solver solver, AwgLaunch L, T *__restrict__ gscratch) {
    extern __shared__ __align__(16) unsigned char awg_lds[];
    T *lds = reinterpret_cast<T *>(awg_lds);

    const i64 bg2 = (i64)blockIdx.x % L.g2;
    const i64 bg0 = (i64)blockIdx.x / L.g2;
    const i64 l2 = threadIdx.x, l1 = threadIdx.y, l0 = threadIdx.z;
    const i64 gid0 = bg0 * L.w0 + l0;
    const i64 gid2 = bg2 * L.w2 + l2;

    const i64 window = solver.window();
    const i64 nw = L.n1 - (window - 1);

    // MemAllocator<Local>: scratch (w0, w2, nw) indexed by local id;
    // MemAllocator<Global>: caller span indexed by global id
    // (src/core/MemorySpace.hpp:14-70).
    T *scr = GLOBAL_SCRATCH ? gscratch + (gid0 * L.ge1 + gid2) * L.ge2 : lds + (l0 * L.w2 + l2) * nw;

    // Lines of this batch only (the reference bounds the loop by start +
    // batch size instead).
    const i64 start0 = L.b0_offset + gid0;
    const i64 stop0 = min(L.n0, L.b0_offset + L.b0_size);
    const i64 start2 = L.b2_offset + gid2;
    const i64 stop2 = min(L.n2, L.b2_offset + L.b2_size);

    for (i64 t0 = 0; t0 < L.trips0; ++t0) {
        const i64 i0 = start0 + t0 * L.g0 * L.w0;
        for (i64 t2 = 0; t2 < L.trips2; ++t2) {
            const i64 i2 = start2 + t2 * L.g2 * L.w2;
            const bool active = (i0 < stop0) && (i2 < stop2);
            const i64 gi0 = L.i0_base + i0;
            Line<T> line{data + (gi0 * L.n1) * L.n2 + i2, L.n2};
            if (active) {
                for (i64 ii1 = l1; ii1 < L.n1; ii1 += L.w1) {
                    const i64 iw = ii1 - (window - 1);
                    if (iw >= 0) scr[iw] = (T)solver(line, gi0, ii1, i2);
                }
            }
            __syncthreads();
            if (active) {
                for (i64 iw = l1; iw < nw; iw += L.w1) line.at(iw) = scr[iw];
            }
            __syncthreads();
        }
    }
}

// NDRange: one work-group of n1 work-items per line, LDS scratch of n1
// elements.  Like the reference's kernel it stores row i1's result at i1
// (no window offset).
template <class T, class Solver>
__global__ void ndrange_kernel(T *__restrict__ data, Solver solver, i64 n1, i64 n2, i64 line_offset) {
    extern __shared__ __align__(16) unsigned char ndr_lds[];
    T *lds = reinterpret_cast<T *>(ndr_lds);
    const i64 line_id = line_offset + (i64)blockIdx.x;
    const i64 i0 = line_id / n2, i2 = line_id % n2;
    const i64 i1 = threadIdx.x;
    Line<T> line{data + (i0 * n1) * n2 + i2, n2};
    lds[i1] = (T)solver(line, i0, i1, i2);
    __syncthreads();
    line.at(i1) = lds[i1];
}

// BasicRange: out of place into a global (n0, n1, n2) scratch, then a copy
// (no window offset, as in the reference).
template <class T, class Solver>
__global__ void basicrange_compute_kernel(const T *__restrict__ data, Solver solver, i64 n0, i64 n1, i64 n2,
                                          T *__restrict__ scratch) {
    const i64 total = n0 * n1 * n2;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const i64 i2 = idx % n2;
        const i64 i1 = (idx / n2) % n1;
        const i64 i0 = idx / (n1 * n2);
        Line<const T> line{data + (i0 * n1) * n2 + i2, n2};
        scratch[idx] = (T)solver(line, i0, i1, i2);
    }
}

template <class T> __global__ void copy_kernel(T *__restrict__ dst, const T *__restrict__ src, i64 n) {
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (i64)gridDim.x * blockDim.x)
        dst[idx] = src[idx];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
constexpr i64 kMaxWgSize = 1024;

inline Status lds_limit(u64 *bytes) {
    int dev = 0;
    BKMA_EXEC_TRY(hip_status(hipGetDevice(&dev), "hipGetDevice"));
    int v = 0;
    BKMA_EXEC_TRY(hip_status(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                             "hipDeviceGetAttribute"));
    *bytes = (u64)v;
    return Status{};
}

// Host mirror of the launch-time recomputation in AdaptiveWg.hpp:66-73:
// WorkGroupDispatch::set_num_work_groups(b0, b2, 1, 1, w0, w2) with the
// dispatch's s0_/s2_ (src/core/bkma_tools.hpp:34-55).
inline Status adaptivewg_shape(const bkma_optim_params &o, i64 b0_size, i64 b2_size, i64 &w0, i64 &w2, i64 &g0,
                               i64 &g2) {
    w0 = std::min<i64>((i64)o.w0, b0_size);
    w2 = std::min<i64>((i64)o.w2, b2_size);
    if (w0 <= 0 || w2 <= 0 || o.w1 == 0)
        return error(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg: work-group sizes must be positive");
    const i64 s0 = (i64)std::max<u64>(o.wg_dispatch.s0, 1), s2 = (i64)std::max<u64>(o.wg_dispatch.s2, 1);
    if (s0 > b0_size || s2 > b2_size)
        return error(BKMA_ERR_INVALID_ARGUMENT,
                     "s0_ > n0 || s2_ > n2. Sequential size cannot be larger than dimension size.");
    if (s0 * w0 > b0_size || s2 * w2 > b2_size)
        return error(BKMA_ERR_INVALID_ARGUMENT,
                     "s0_*w0 > n0 || s2_*w2 > n2. A single work-group cannot process more than dimension size.");
    g0 = b0_size / s0 / w0;
    g2 = b2_size / s2 / w2;
    return Status{};
}

// submit_kernels<MemType, Solver, AdaptiveWg> in the host batch loop of
// bkma_run.  `i0_base` offsets the launch by whole slabs (n0 is then the
// number of slabs from there).
template <class T, class Solver>
Status run_adaptivewg(T *data, const Solver &solver, i64 n0, i64 n1, i64 n2, const bkma_optim_params &o,
                      const bkma_span3d &scratch, hipStream_t stream, i64 i0_base = 0) {
    const i64 window = solver.window();
    if (window < 1 || window > n1) return error(BKMA_ERR_INVALID_ARGUMENT, "solver window larger than n1");
    const i64 nw = n1 - (window - 1);
    const bool global = o.mem_space == BKMA_MEM_GLOBAL;
    if (o.mem_space != BKMA_MEM_LOCAL && !global) return error(BKMA_ERR_INVALID_ARGUMENT, "Unknown MemorySpace");
    const auto &d0 = o.dispatch_d0, &d2 = o.dispatch_d2;
    if (d0.n_batch == 0 || d2.n_batch == 0)
        return error(BKMA_ERR_INVALID_ARGUMENT, "BatchConfig1D with zero batches");

    u64 lds_max = 0;
    BKMA_EXEC_TRY(lds_limit(&lds_max));

    // The reference waits on the previous batch's event before each
    // submission; launches on one stream are already ordered.
    for (u64 ib0 = 0; ib0 < d0.n_batch; ++ib0) {
        const bool last0 = ib0 == d0.n_batch - 1;
        const i64 off0 = (i64)(d0.batch_size * ib0);
        const i64 bs0 = (i64)(last0 ? d0.last_batch_size : d0.batch_size);
        for (u64 ib2 = 0; ib2 < d2.n_batch; ++ib2) {
            const bool last2 = ib2 == d2.n_batch - 1;
            const i64 off2 = (i64)(d2.batch_size * ib2);
            const i64 bs2 = (i64)(last2 ? d2.last_batch_size : d2.batch_size);
            if (bs0 <= 0 || bs2 <= 0) continue;

            i64 w0, w2, g0, g2;
            BKMA_EXEC_TRY(adaptivewg_shape(o, bs0, bs2, w0, w2, g0, g2));
            const i64 w1 = (i64)o.w1;
            if (w0 * w1 * w2 > kMaxWgSize)
                return error(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg: work-group size w0*w1*w2 exceeds 1024");

            AwgLaunch L;
            L.n0 = n0;
            L.n1 = n1;
            L.n2 = n2;
            L.b0_size = bs0;
            L.b0_offset = off0;
            L.b2_size = bs2;
            L.b2_offset = off2;
            L.w0 = w0;
            L.w1 = w1;
            L.w2 = w2;
            L.g0 = g0;
            L.g2 = g2;
            L.trips0 = (i64)div_up((u64)bs0, (u64)(g0 * w0));
            L.trips2 = (i64)div_up((u64)bs2, (u64)(g2 * w2));
            L.ge1 = 0;
            L.ge2 = 0;
            L.i0_base = i0_base;

            const dim3 block((unsigned)w2, (unsigned)w1, (unsigned)w0);
            const u64 nblocks = (u64)g0 * (u64)g2;
            if (nblocks == 0 || nblocks > 0x7fffffffULL)
                return error(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg: work-group count out of range");

            if (!global) {
                const u64 lds_bytes = (u64)w0 * (u64)w2 * (u64)nw * sizeof(T);
                if (lds_bytes > lds_max)
                    return error(BKMA_ERR_INVALID_ARGUMENT,
                                 "AdaptiveWg/Local: scratch (w0, w2, nw) of " + std::to_string(lds_bytes) +
                                     " bytes exceeds the work-group LDS limit of " + std::to_string(lds_max));
                auto kern = adaptivewg_kernel<T, Solver, false>;
                BKMA_EXEC_TRY(hip_status(hipFuncSetAttribute((const void *)kern,
                                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                                             (int)lds_bytes),
                                         "hipFuncSetAttribute"));
                hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), block, (size_t)lds_bytes, stream, data, solver, L,
                                   (T *)nullptr);
            } else {
                // The global scratch must cover (g0*w0, g2*w2, nw).
                if (scratch.data == nullptr)
                    return error(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg/Global: a global scratch span is required");
                if ((i64)scratch.e0 < g0 * w0 || (i64)scratch.e1 < g2 * w2 || (i64)scratch.e2 < nw)
                    return error(BKMA_ERR_INVALID_ARGUMENT,
                                 "AdaptiveWg/Global: scratch extents must cover (g0*w0, g2*w2, n1-window+1) = (" +
                                     std::to_string(g0 * w0) + ", " + std::to_string(g2 * w2) + ", " +
                                     std::to_string(nw) + ")");
                L.ge1 = (i64)scratch.e1;
                L.ge2 = (i64)scratch.e2;
                hipLaunchKernelGGL((adaptivewg_kernel<T, Solver, true>), dim3((unsigned)nblocks), block, 0, stream,
                                   data, solver, L, reinterpret_cast<T *>(scratch.data));
            }
            BKMA_EXEC_TRY(launch_status("adaptivewg_kernel"));
        }
    }
    return Status{};
}

template <class T, class Solver>
Status run_ndrange(T *data, const Solver &solver, i64 n0, i64 n1, i64 n2, hipStream_t stream) {
    if (n1 > kMaxWgSize)
        return error(BKMA_ERR_INVALID_ARGUMENT,
                     "NDRange: n1 = " + std::to_string(n1) + " exceeds the maximum work-group size of 1024");
    u64 lds_max = 0;
    BKMA_EXEC_TRY(lds_limit(&lds_max));
    const u64 lds_bytes = (u64)n1 * sizeof(T);
    if (lds_bytes > lds_max) return error(BKMA_ERR_INVALID_ARGUMENT, "NDRange: line does not fit in LDS");
    const i64 lines = n0 * n2;
    const i64 max_blocks = 0x7fffffffLL / std::max<i64>(n1, 1);
    for (i64 off = 0; off < lines; off += max_blocks) {
        const i64 nb = std::min(max_blocks, lines - off);
        hipLaunchKernelGGL((ndrange_kernel<T, Solver>), dim3((unsigned)nb), dim3((unsigned)n1), (size_t)lds_bytes,
                           stream, data, solver, n1, n2, off);
        BKMA_EXEC_TRY(launch_status("ndrange_kernel"));
    }
    return Status{};
}

template <class T, class Solver>
Status run_basicrange(T *data, const Solver &solver, i64 n0, i64 n1, i64 n2, const bkma_span3d &scratch,
                      hipStream_t stream) {
    if (scratch.data == nullptr)
        return error(BKMA_ERR_INVALID_ARGUMENT, "BasicRange: a (n0, n1, n2) global scratch span is required");
    if ((i64)scratch.e0 != n0 || (i64)scratch.e1 != n1 || (i64)scratch.e2 != n2)
        return error(BKMA_ERR_INVALID_ARGUMENT, "BasicRange: scratch extents must equal (n0, n1, n2)");
    T *scr = reinterpret_cast<T *>(scratch.data);
    const i64 total = n0 * n1 * n2;
    const unsigned block = 256;
    const u64 grid = std::min<u64>(div_up((u64)total, block), 1u << 20);
    hipLaunchKernelGGL((basicrange_compute_kernel<T, Solver>), dim3((unsigned)grid), dim3(block), 0, stream,
                       (const T *)data, solver, n0, n1, n2, scr);
    BKMA_EXEC_TRY(launch_status("basicrange_compute_kernel"));
    hipLaunchKernelGGL(copy_kernel<T>, dim3((unsigned)grid), dim3(block), 0, stream, data, (const T *)scr, total);
    return launch_status("copy_kernel");
}

// bkma_run<Solver, Impl> (src/core/bkma_run.hpp:10-62) on an [n0][n1][n2]
// array of T, enqueued on `stream`.
template <class T, class Solver>
Status run(bkma_impl impl, T *data, const Solver &solver, i64 n0, i64 n1, i64 n2, const bkma_optim_params *optim,
           const bkma_span3d &scratch, hipStream_t stream) {
    if (n0 <= 0 || n1 <= 0 || n2 <= 0) return error(BKMA_ERR_INVALID_ARGUMENT, "array extents must be positive");
    if (n1 > 0x7fffffffLL || n0 > 0x7fffffffLL)
        return error(BKMA_ERR_UNSUPPORTED, "n0 and n1 must fit in a 32-bit int (the solver's index type)");
    switch (impl) {
        case BKMA_IMPL_ADAPTIVE_WG:
            if (!optim) return error(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg needs BkmaOptimParams");
            return run_adaptivewg(data, solver, n0, n1, n2, *optim, scratch, stream);
        case BKMA_IMPL_ND_RANGE:
            return run_ndrange(data, solver, n0, n1, n2, stream);
        case BKMA_IMPL_BASIC_RANGE:
            return run_basicrange(data, solver, n0, n1, n2, scratch, stream);
    }
    return error(BKMA_ERR_RUNTIME, "not a valid implementation. Should be: {BasicRange, NDRange, AdaptiveWg}");
}

}  // namespace exec
}  // namespace bkma
