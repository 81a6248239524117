// Ragged batches of independent advection problems in one persistent launch.
//
// Each problem ([n0][n1][n2], its own dt and velocity grid) is cut into work
// items of the LINE_LDS kind.  A problem whose slab fits the LDS budget is
// one item per slab (whole line, periodic wrap in LDS); a longer one is cut
// into segments, and an item is a chunk of consecutive segments of one slab
// that a work-group walks in order (line_chunk_unit): only the halo rows
// around the chunk come from a snapshot buffer filled by batch_halo_kernel
// just before.  One pass of the batch advances every problem by ns steps
// (steps fused in LDS, intermediate steps in FP64 as in the LINE_LDS plans:
// whole lines by line_steps_unit, chunks by line_chunk_steps_unit with halos
// of ns steps):
//   batch_halo_kernel  (only if some problem is segmented)
//   adv_batch_kernel   (persistent: work-groups pull items from a counter)
// and runs of steps are captured into hipGraphs of up to 32 steps and replayed.
//
// The work counter is never reset: every execution of adv_batch_kernel hands
// out exactly nitems valid tickets plus one terminating ticket per work-group,
// so ticket mod (nitems + grid) is the item index of the current execution.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "adv_lds.hpp"
#include "common.hpp"

namespace bkma {

struct BatchProblemDev {
    void *data;
    void *halo;  // segmented problems: snapshot buffer per (slab, chunk) (nullptr otherwise)
    AdvDev p;
    i64 seg_rows, nseg, hb_max, ha_max;
    i64 chunk_segs, nchunk;  // segmented problems: segments walked in order by one item
};

struct BatchItem {
    int32_t prob;
    int32_t chunk;  // segmented problems: chunk of chunk_segs segments (else 0)
    i64 i0;
};

// FUSED: passes of ns >= 2 steps (256-thread work-groups); a kernel of its
// own, so that the one-step kernel keeps its registers and occupancy.
template <class T, int NT, bool FUSED>
__global__ __launch_bounds__(NT) void adv_batch_kernel(const BatchProblemDev *__restrict__ probs,
                                                        const BatchItem *__restrict__ items, i64 nitems,
                                                        unsigned long long *counter, int ns) {
    static_assert(!FUSED || NT == 256, "fused passes run in 256-thread work-groups");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ i64 s_item;
    const i64 modulus = nitems + (i64)gridDim.x;
    for (;;) {
        if (threadIdx.x == 0) s_item = (i64)(atomicAdd(counter, 1ull) % (unsigned long long)modulus);
        __syncthreads();
        const i64 it = s_item;
        __syncthreads();  // everyone has read s_item before it is overwritten
        if (it >= nitems) break;  // work-group-uniform exit
        const BatchItem item = items[it];
        const BatchProblemDev &pr = probs[item.prob];
        const i64 n1 = pr.p.n1, n2 = pr.p.n2;
        T *slab = reinterpret_cast<T *>(pr.data) + item.i0 * n1 * n2;
        T *st = reinterpret_cast<T *>(lds_raw);
        if (pr.nseg > 1) {
            const i64 chunk_rows = pr.seg_rows * pr.chunk_segs;
            const i64 c0 = (i64)item.chunk * chunk_rows;
            const i64 c1 = min(n1, c0 + chunk_rows);
            const T *hsnap = reinterpret_cast<const T *>(pr.halo) +
                             (item.i0 * pr.nchunk + item.chunk) * ns * (pr.hb_max + pr.ha_max) * n2;
            if constexpr (FUSED) {
                // staged rows, then two FP64 buffers of as many rows (16-byte aligned)
                const i64 rows = ns * (pr.hb_max + pr.ha_max) + pr.seg_rows;
                double *w0 = reinterpret_cast<double *>(lds_raw + ((rows * n2 * (i64)sizeof(T) + 15) & ~(i64)15));
                line_chunk_steps_unit<T>(slab, hsnap, pr.p, item.i0, c0, c1, pr.seg_rows, pr.hb_max, pr.ha_max, ns,
                                         st, w0, w0 + rows * n2);
            } else {
                line_chunk_unit<T>(slab, hsnap, pr.p, item.i0, c0, c1, pr.seg_rows, pr.hb_max, pr.ha_max, st);
            }
        } else if constexpr (FUSED) {
            double *w0 = reinterpret_cast<double *>(lds_raw + ((n1 * n2 * (i64)sizeof(T) + 15) & ~(i64)15));
            line_steps_unit<T, line_feet_rows<NT>()>(slab, pr.p, item.i0, 1, ns, st, w0, w0 + n1 * n2);
        } else {
            line_lds_unit<T, false>(slab, nullptr, pr.p, item.i0, 0, n1, 0, 0, st);
        }
        __syncthreads();  // LDS is reused by the next item
    }
}

// Halo snapshot of the segmented items (one work-group per item): the hb
// rows before and the ha rows after the item's chunk.
template <class T>
__global__ void batch_halo_kernel(const BatchProblemDev *__restrict__ probs, const BatchItem *__restrict__ items,
                                  int ns) {
    const BatchItem item = items[blockIdx.x];
    const BatchProblemDev &pr = probs[item.prob];
    const i64 n1 = pr.p.n1, n2 = pr.p.n2;
    const i64 r0 = (i64)item.chunk * pr.seg_rows * pr.chunk_segs;
    const i64 r1 = min(n1, r0 + pr.seg_rows * pr.chunk_segs);
    int hb, ha;
    slab_halo(pr.p, item.i0, pr.hb_max, pr.ha_max, hb, ha);
    hb *= ns;  // ns fused steps reach ns times as far
    ha *= ns;
    const T *slab = reinterpret_cast<const T *>(pr.data) + item.i0 * n1 * n2;
    T *dst = reinterpret_cast<T *>(pr.halo) + (item.i0 * pr.nchunk + item.chunk) * ns * (pr.hb_max + pr.ha_max) * n2;
    const i64 total = (i64)(hb + ha) * n2;
    for (i64 e = threadIdx.x; e < total; e += blockDim.x) {
        const i64 hr = e / n2, c = e - hr * n2;
        i64 row = hr < hb ? r0 - hb + hr : r1 + (hr - hb);
        row %= n1;
        if (row < 0) row += n1;
        dst[e] = slab[row * n2 + c];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

constexpr i64 kBatchChunkSegs = 16;  // segments one item walks in order

void shift_bounds(const bkma_adv_params &p, int *dlo, int *dhi) {
    const double v0 = p.minRealVx, v1 = p.minRealVx + (double)(p.n0 - 1) * p.dvx;
    const double s0 = v0 * p.dt * p.inv_dx, s1 = v1 * p.dt * p.inv_dx;
    *dlo = (int)std::floor(std::min(s0, s1) - 1e-6);
    *dhi = (int)std::floor(std::max(s0, s1) + 1e-6) + 1;
}

}  // namespace

struct AdvBatch {
    bkma_dtype dtype = BKMA_F32;
    std::vector<bkma_adv_problem> problems;
    std::vector<BatchProblemDev> host_probs;
    std::vector<BatchItem> host_items, host_seg_items;
    BatchProblemDev *d_probs = nullptr;
    BatchItem *d_items = nullptr, *d_seg_items = nullptr;
    unsigned long long *d_counter = nullptr;
    std::vector<void *> halos;
    u64 lds_budget = 0;
    int threads = 256;  // per work-group: 256 or 1024 (the autotuner's choice)
    int ns = 1;         // steps per pass (fused steps need 256-thread work-groups)
    unsigned grid = 0;
    hipStream_t stream = nullptr;  // internal stream (graph capture and replay)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::map<u64, hipGraphExec_t> graphs;  // by step count
    std::vector<bkma_adv_plan *> plans;    // per-problem tuned plans (loop mode)
    u64 cells = 0;

    ~AdvBatch() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        for (auto *pl : plans) bkma_adv_plan_destroy(pl);
        for (void *h : halos) (void)hipFree(h);
        if (d_probs) (void)hipFree(d_probs);
        if (d_items) (void)hipFree(d_items);
        if (d_seg_items) (void)hipFree(d_seg_items);
        if (d_counter) (void)hipFree(d_counter);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// One pass of ns <= b.ns steps (the items, halo buffers and LDS budget are
// sized for b.ns, and hold for fewer steps).
template <class T> static bkma_status batch_pass(const AdvBatch &b, hipStream_t s, int ns) {
    if (!b.host_seg_items.empty()) {
        hipLaunchKernelGGL(batch_halo_kernel<T>, dim3((unsigned)b.host_seg_items.size()), dim3(256), 0, s,
                           (const BatchProblemDev *)b.d_probs, (const BatchItem *)b.d_seg_items, ns);
        BKMA_TRY(check_launch("batch_halo_kernel"));
    }
    // a pass of one step (a run's remainder) runs the one-step kernel: the
    // fused chunk unit needs ns >= 2
    auto k = ns > 1 ? adv_batch_kernel<T, 256, true>
                    : (b.threads == 1024 ? adv_batch_kernel<T, 1024, false> : adv_batch_kernel<T, 256, false>);
    hipLaunchKernelGGL(k, dim3(b.grid), dim3(b.threads), (size_t)b.lds_budget, s, (const BatchProblemDev *)b.d_probs,
                       (const BatchItem *)b.d_items, (i64)b.host_items.size(), b.d_counter, ns);
    return check_launch("adv_batch_kernel");
}

// nsteps steps: passes of b.ns steps, then one pass of the remainder.
static bkma_status batch_steps(const AdvBatch &b, u64 nsteps, hipStream_t s) {
    for (u64 t = 0; t < nsteps;) {
        const int k = (int)std::min<u64>((u64)b.ns, nsteps - t);
        BKMA_TRY(b.dtype == BKMA_F64 ? batch_pass<double>(b, s, k) : batch_pass<float>(b, s, k));
        t += (u64)k;
    }
    return BKMA_OK;
}

// Build the item lists for an LDS budget (bytes per work-group), work-groups
// of `threads` threads and ns steps per pass.  With ns > 1 a staged row takes
// two FP64 rows besides its own, and a chunk's halos are ns steps wide.
static bkma_status build_items(AdvBatch &b, u64 budget, int threads, int ns) {
    if (ns < 1 || (ns > 1 && threads > 256))  // the kernel fuses steps in 256-thread work-groups only
        return fail(BKMA_ERR_UNSUPPORTED, "advection batch: fused steps need 256-thread work-groups");
    const u64 esz = b.dtype == BKMA_F64 ? 8 : 4;
    const u64 row_esz = esz + (ns > 1 ? 16 : 0);  // LDS bytes per staged cell
    for (void *h : b.halos) (void)hipFree(h);
    b.halos.clear();
    b.host_probs.clear();
    b.host_items.clear();
    b.host_seg_items.clear();
    for (size_t j = 0; j < b.problems.size(); ++j) {
        const bkma_adv_params &p = b.problems[j].params;
        BatchProblemDev d{};
        d.data = b.problems[j].data;
        d.p = make_adv_dev(p);
        const u64 slab_bytes = p.n1 * p.n2 * row_esz + 16;
        if (slab_bytes <= budget) {
            d.seg_rows = (i64)p.n1;
            d.nseg = 1;
            d.chunk_segs = d.nchunk = 1;
        } else {
            int dlo, dhi;
            shift_bounds(p, &dlo, &dhi);
            d.hb_max = std::max(0, dhi + 2);
            d.ha_max = std::max(0, 3 - dlo);
            const i64 rows = (i64)((budget - 16) / (p.n2 * row_esz)) - (i64)ns * (d.hb_max + d.ha_max);
            if (rows < 16)
                return fail(BKMA_ERR_UNSUPPORTED, "advection batch: problem " + std::to_string(j) +
                                                      " has rows too wide or a shift too large for the LDS budget");
            d.seg_rows = rows & ~(i64)3;
            d.nseg = (i64)ceil_div(p.n1, (u64)d.seg_rows);
            // chain up to kBatchChunkSegs segments per item (needs segments at
            // least as long as the halo before them)
            d.chunk_segs = d.seg_rows >= (i64)ns * d.hb_max ? std::min<i64>(d.nseg, kBatchChunkSegs) : 1;
            d.nchunk = (i64)ceil_div((u64)d.nseg, (u64)d.chunk_segs);
            void *h = nullptr;
            const u64 hb = p.n0 * (u64)d.nchunk * (u64)ns * (u64)(d.hb_max + d.ha_max) * p.n2 * esz;
            BKMA_HIP_TRY(hipMalloc(&h, std::max<u64>(hb, 16)));
            b.halos.push_back(h);
            d.halo = h;
        }
        b.host_probs.push_back(d);
        for (i64 i0 = 0; i0 < (i64)p.n0; ++i0)
            for (i64 ch = 0; ch < d.nchunk; ++ch) {
                BatchItem it{(int32_t)j, (int32_t)ch, i0};
                b.host_items.push_back(it);
                if (d.nseg > 1) b.host_seg_items.push_back(it);
            }
    }
    // Largest items first: the persistent work-groups then finish together.
    std::stable_sort(b.host_items.begin(), b.host_items.end(), [&](const BatchItem &x, const BatchItem &y) {
        auto cost = [&](const BatchItem &i) {
            const auto &d = b.host_probs[i.prob];
            return (double)std::min<i64>(d.seg_rows * d.chunk_segs, d.p.n1) * (double)d.p.n2;
        };
        return cost(x) > cost(y);
    });
    if (b.d_probs) (void)hipFree(b.d_probs);
    if (b.d_items) (void)hipFree(b.d_items);
    if (b.d_seg_items) (void)hipFree(b.d_seg_items);
    b.d_probs = nullptr;
    b.d_items = b.d_seg_items = nullptr;
    BKMA_HIP_TRY(hipMalloc((void **)&b.d_probs, sizeof(BatchProblemDev) * b.host_probs.size()));
    BKMA_HIP_TRY(hipMemcpy(b.d_probs, b.host_probs.data(), sizeof(BatchProblemDev) * b.host_probs.size(),
                           hipMemcpyHostToDevice));
    BKMA_HIP_TRY(hipMalloc((void **)&b.d_items, sizeof(BatchItem) * std::max<size_t>(1, b.host_items.size())));
    BKMA_HIP_TRY(hipMemcpy(b.d_items, b.host_items.data(), sizeof(BatchItem) * b.host_items.size(),
                           hipMemcpyHostToDevice));
    if (!b.host_seg_items.empty()) {
        BKMA_HIP_TRY(hipMalloc((void **)&b.d_seg_items, sizeof(BatchItem) * b.host_seg_items.size()));
        BKMA_HIP_TRY(hipMemcpy(b.d_seg_items, b.host_seg_items.data(), sizeof(BatchItem) * b.host_seg_items.size(),
                               hipMemcpyHostToDevice));
    }
    // Persistent grid: every work-group that can be resident at this LDS size.
    b.lds_budget = budget;
    b.threads = threads;
    b.ns = ns;
    int dev = 0, cus = 0, blocks = 0;
    BKMA_HIP_TRY(hipGetDevice(&dev));
    BKMA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void *k = b.dtype == BKMA_F64
                        ? (ns > 1 ? (const void *)adv_batch_kernel<double, 256, true>
                                  : (threads == 1024 ? (const void *)adv_batch_kernel<double, 1024, false>
                                                     : (const void *)adv_batch_kernel<double, 256, false>))
                        : (ns > 1 ? (const void *)adv_batch_kernel<float, 256, true>
                                  : (threads == 1024 ? (const void *)adv_batch_kernel<float, 1024, false>
                                                     : (const void *)adv_batch_kernel<float, 256, false>));
    BKMA_HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)budget));
    BKMA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, threads, (size_t)budget));
    b.grid = (unsigned)std::max<i64>(1, std::min<i64>((i64)blocks * cus, (i64)b.host_items.size()));
    // Fresh counter (the modulus changed with the item count / grid).
    BKMA_HIP_TRY(hipMemset(b.d_counter, 0, sizeof(unsigned long long)));
    for (auto &g : b.graphs) (void)hipGraphExecDestroy(g.second);
    b.graphs.clear();
    return BKMA_OK;
}

}  // namespace bkma

using namespace bkma;

struct bkma_adv_batch : AdvBatch {};

// Validation, counter, stream and events of a new batch (no items yet).
static bkma_status batch_setup(const bkma_adv_problem *problems, int32_t nproblems, bkma_dtype dtype,
                               bkma_adv_batch **out) {
    if (!problems || !out || nproblems <= 0) return fail(BKMA_ERR_INVALID_ARGUMENT, "advection batch: no problems");
    *out = nullptr;
    if (dtype != BKMA_F32 && dtype != BKMA_F64) return fail(BKMA_ERR_UNSUPPORTED, "advection batch: F32 or F64");
    BKMA_TRY(require_device());
    auto *b = new (std::nothrow) bkma_adv_batch();
    if (!b) return fail(BKMA_ERR_OUT_OF_MEMORY, "batch allocation");
    b->dtype = dtype;
    for (int32_t j = 0; j < nproblems; ++j) {
        const bkma_adv_params &p = problems[j].params;
        const double smax = std::fabs(p.dt * p.inv_dx) * std::max(std::fabs(p.minRealVx), std::fabs(p.maxRealVx));
        if (!problems[j].data || reinterpret_cast<uintptr_t>(problems[j].data) % 16 != 0 || p.n0 == 0 || p.n1 == 0 ||
            p.n2 == 0 || p.n0 > 0x7fffffffULL ||
            p.n1 > 0x7fffffffULL || !std::isfinite(p.dt) || !(p.realWidthX > 0) || !(smax < (double)p.n1 / 2 - 8)) {
            delete b;
            return fail(BKMA_ERR_INVALID_ARGUMENT,
                        "advection batch: problem " + std::to_string(j) + " is invalid (or its data is not 16-byte aligned)");
        }
        b->problems.push_back(problems[j]);
        b->cells += p.n0 * p.n1 * p.n2;
    }
    auto bail = [&](bkma_status s) {
        delete b;
        return s;
    };
    if (hipMalloc((void **)&b->d_counter, sizeof(unsigned long long)) != hipSuccess)
        return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "batch counter"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess)
        return bail(fail(BKMA_ERR_DEVICE, "batch stream/event creation"));
    *out = b;
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_batch_create_with(const bkma_adv_problem *problems, int32_t nproblems,
                                                  bkma_dtype dtype, uint64_t lds_budget, int32_t threads,
                                                  int32_t steps_per_pass, void *stream, bkma_adv_batch **out) {
    (void)stream;
    u64 lds_max = 0;
    BKMA_TRY(lds_limit_bytes(&lds_max));
    if ((threads != 256 && threads != 1024) || steps_per_pass < 1 || steps_per_pass > 4 || lds_budget < 1024 ||
        lds_budget > lds_max - 1024)
        return fail(BKMA_ERR_INVALID_ARGUMENT, "advection batch: threads 256 or 1024, 1 to 4 steps per pass, "
                                               "an LDS budget from 1 KB to the device limit less 1 KB");
    bkma_adv_batch *b = nullptr;
    BKMA_TRY(batch_setup(problems, nproblems, dtype, &b));
    const bkma_status st = build_items(*b, lds_budget, threads, steps_per_pass);
    if (st != BKMA_OK) {
        delete b;
        return st;
    }
    *out = b;
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_batch_create(const bkma_adv_problem *problems, int32_t nproblems, bkma_dtype dtype,
                                             int32_t autotune, void *stream, bkma_adv_batch **out) {
    u64 lds_max = 0;
    BKMA_TRY(lds_limit_bytes(&lds_max));
    bkma_adv_batch *b = nullptr;
    BKMA_TRY(batch_setup(problems, nproblems, dtype, &b));
    bkma_status st = BKMA_OK;
    auto bail = [&](bkma_status s) {
        delete b;
        return s;
    };

    // The timing runs below use the internal stream: order them after the
    // caller's pending work on `stream` (its writes to the problems' arrays).
    if (autotune) {
        if (hipEventRecord(b->ev_in, as_stream(stream)) != hipSuccess ||
            hipStreamWaitEvent(b->stream, b->ev_in, 0) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "batch: ordering after the caller's stream"));
    }

    // Dynamic LDS budgets per work-group (the kernel adds a few static bytes),
    // work-group sizes and steps per pass.  Without autotuning: 64 KB, 256
    // threads, one step per pass.
    const u64 cap = lds_max - 1024;
    std::vector<u64> budgets = {std::min<u64>(65536, cap)};
    std::vector<int> passes = {1};
    if (autotune) {
        budgets = {32768, 65536, 98304, cap};
        passes = {1, 2, 3, 4};
    }
    u64 best_budget = 0;
    int best_threads = 256, best_ns = 1;
    float best_ms = 1e30f;
    bool done = false;
    for (u64 bud : budgets) {
        for (int thr : {256, 1024}) {
            for (int ns : passes) {
                if (done || bud > lds_max || (ns > 1 && thr > 256)) break;
                if (build_items(*b, bud, thr, ns) != BKMA_OK) continue;
                if (!autotune) {
                    best_budget = bud;
                    done = true;
                    break;
                }
                // The timing runs update the problems in place: the autotuner is
                // meant to run before the data is initialised (or on scratch data).
                hipEvent_t e0, e1;
                (void)hipEventCreate(&e0);
                (void)hipEventCreate(&e1);
                st = batch_steps(*b, (u64)ns, b->stream);
                (void)hipEventRecord(e0, b->stream);
                if (st == BKMA_OK) st = batch_steps(*b, 3 * (u64)ns, b->stream);
                (void)hipEventRecord(e1, b->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                ms /= (float)ns;  // per 3 steps
                if (st == BKMA_OK && ms < best_ms) {
                    best_ms = ms;
                    best_budget = bud;
                    best_threads = thr;
                    best_ns = ns;
                }
                st = BKMA_OK;
            }
        }
    }
    if (best_budget == 0) return bail(fail(BKMA_ERR_UNSUPPORTED, "advection batch: no LDS budget covers the problems"));
    st = build_items(*b, best_budget, best_threads, best_ns);
    if (st != BKMA_OK) return bail(st);
    if (autotune) {  // later work on the caller's stream sees the timing runs' updates
        if (hipEventRecord(b->ev_out, b->stream) != hipSuccess ||
            hipStreamWaitEvent(as_stream(stream), b->ev_out, 0) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "batch: ordering the caller's stream after the timing runs"));
    }
    *out = b;
    return BKMA_OK;
}

// Time steps per captured graph: a run of n steps replays the graph of
// kGraphSteps steps n / kGraphSteps times, then the graph of the remainder.
constexpr u64 kGraphSteps = 32;

// The instantiated graph of `nsteps` (<= kGraphSteps) batch steps, captured
// on the batch's stream the first time it is asked for.
static bkma_status batch_graph(bkma_adv_batch &b, u64 nsteps, hipGraphExec_t *out) {
    auto it = b.graphs.find(nsteps);
    if (it == b.graphs.end()) {
        hipGraph_t g = nullptr;
        BKMA_HIP_TRY(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
        bkma_status st = batch_steps(b, nsteps, b.stream);
        hipError_t ce = hipStreamEndCapture(b.stream, &g);
        if (st != BKMA_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
        hipGraphExec_t ge = nullptr;
        hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate");
        it = b.graphs.emplace(nsteps, ge).first;
    }
    *out = it->second;
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_batch_run(bkma_adv_batch *b, uint64_t nsteps, int32_t use_graph, void *stream) {
    if (!b) return fail(BKMA_ERR_INVALID_ARGUMENT, "null batch");
    if (nsteps == 0) return BKMA_OK;
    hipStream_t cs = as_stream(stream);
    // Order the internal stream after the caller's stream, and back.
    BKMA_HIP_TRY(hipEventRecord(b->ev_in, cs));
    BKMA_HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    if (use_graph) {
        // Graphs of kGraphSteps steps and of the remainder: at most
        // kGraphSteps graphs per batch, each captured on first use.
        const u64 full = nsteps / kGraphSteps, rem = nsteps % kGraphSteps;
        hipGraphExec_t gfull = nullptr, grem = nullptr;
        if (full) BKMA_TRY(batch_graph(*b, kGraphSteps, &gfull));
        if (rem) BKMA_TRY(batch_graph(*b, rem, &grem));
        for (u64 i = 0; i < full; ++i) BKMA_HIP_TRY(hipGraphLaunch(gfull, b->stream));
        if (rem) BKMA_HIP_TRY(hipGraphLaunch(grem, b->stream));
    } else {
        BKMA_TRY(batch_steps(*b, nsteps, b->stream));
    }
    BKMA_HIP_TRY(hipEventRecord(b->ev_out, b->stream));
    BKMA_HIP_TRY(hipStreamWaitEvent(cs, b->ev_out, 0));
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_batch_run_loop(bkma_adv_batch *b, uint64_t nsteps, int32_t mode, void *stream) {
    if (!b) return fail(BKMA_ERR_INVALID_ARGUMENT, "null batch");
    hipStream_t cs = as_stream(stream);
    if (mode == 0) {
        // The reference's way: per step, per problem, bkma_run<AdvectionSolver,
        // AdaptiveWg> with create_optim_params (pref_wg_size 256).
        std::vector<bkma_optim_params> opts(b->problems.size());
        const u64 esz = b->dtype == BKMA_F64 ? 8 : 4;
        for (size_t j = 0; j < b->problems.size(); ++j) {
            const auto &p = b->problems[j].params;
            BKMA_TRY(bkma_create_optim_params(p.n0, p.n1, p.n2, 256, 1, 1, esz, 0, &opts[j]));
        }
        for (u64 t = 0; t < nsteps; ++t)
            for (size_t j = 0; j < b->problems.size(); ++j)
                BKMA_TRY(bkma_adv_run(BKMA_IMPL_ADAPTIVE_WG, b->dtype, b->problems[j].data, &b->problems[j].params,
                                      &opts[j], bkma_span3d{nullptr, 0, 0, 0}, stream));
        return BKMA_OK;
    }
    if (mode == 1) {
        // One tuned plan per problem, launched problem after problem.
        if (b->plans.empty()) {
            for (auto &pr : b->problems) {
                bkma_adv_plan *pl = nullptr;
                BKMA_TRY(bkma_adv_plan_create(&pr.params, b->dtype, 0, stream, &pl));
                b->plans.push_back(pl);
            }
        }
        for (u64 t = 0; t < nsteps; ++t)
            for (size_t j = 0; j < b->problems.size(); ++j)
                BKMA_TRY(bkma_adv_plan_run(b->plans[j], b->problems[j].data, 1, stream));
        (void)cs;
        return BKMA_OK;
    }
    return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_adv_batch_run_loop: mode must be 0 or 1");
}

extern "C" bkma_status bkma_adv_batch_describe(const bkma_adv_batch *b, uint64_t *cells, uint64_t *items,
                                               uint64_t *segmented_items, uint64_t *lds_budget, uint32_t *grid) {
    if (!b) return fail(BKMA_ERR_INVALID_ARGUMENT, "null batch");
    if (cells) *cells = b->cells;
    if (items) *items = b->host_items.size();
    if (segmented_items) *segmented_items = b->host_seg_items.size();
    if (lds_budget) *lds_budget = b->lds_budget;
    if (grid) *grid = b->grid;
    return BKMA_OK;
}

extern "C" int32_t bkma_adv_batch_steps_per_pass(const bkma_adv_batch *b) { return b ? b->ns : 0; }

extern "C" void bkma_adv_batch_destroy(bkma_adv_batch *b) {
    if (b) (void)hipStreamSynchronize(b->stream);
    delete b;
}
