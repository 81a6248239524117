/*
 * bkma_amd.h -- C ABI of the MI355X (gfx950) BKMA library.
 *
 * "BKMA" = batched kernels with memory allocations: in-place line operators
 * over a row-major [n0][n1][n2] array, each (i0, :, i2) line updated by a
 * solver functor with a scratch (LDS or global) holding the new values.
 *
 * Conventions
 *  - Every pointer to array data is DEVICE memory allocated by the caller
 *    (hipMalloc, torch tensors, ...).  Nothing here takes ownership.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Work is enqueued asynchronously; the caller synchronises.  Entry points
 *    that return host values (validators, norms) synchronise `stream`.
 *  - Errors are status codes; bkma_last_error() returns the message of the
 *    last failing call on the calling thread.  The C++ facade rethrows them
 *    with the reference's exception types and texts.
 *  - Not thread-safe per stream: calls that share a stream must not overlap.
 *  - Compute paths never fall back to the CPU: with no usable gfx950 device
 *    they return BKMA_ERR_DEVICE.
 */
#ifndef BKMA_AMD_H
#define BKMA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BKMA_AMD_ABI_VERSION 1

typedef enum bkma_status {
    BKMA_OK = 0,
    BKMA_ERR_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument */
    BKMA_ERR_RUNTIME = 2,          /* reference: std::runtime_error */
    BKMA_ERR_DEVICE = 3,           /* HIP runtime failure (reference: sycl::exception) */
    BKMA_ERR_UNSUPPORTED = 4,      /* shape/dtype combination this library does not run */
    BKMA_ERR_OUT_OF_MEMORY = 5
} bkma_status;

/* BkmaImpl, src/core/bkma_tools.hpp:9-13 */
typedef enum bkma_impl {
    BKMA_IMPL_BASIC_RANGE = 0,
    BKMA_IMPL_ND_RANGE = 1,
    BKMA_IMPL_ADAPTIVE_WG = 2
} bkma_impl;

/* MemorySpace, src/core/MemorySpace.hpp:12 */
typedef enum bkma_memspace { BKMA_MEM_LOCAL = 0, BKMA_MEM_GLOBAL = 1 } bkma_memspace;

/* Storage type of a device array. */
typedef enum bkma_dtype {
    BKMA_F64 = 0,
    BKMA_F32 = 1,
    BKMA_BF16 = 2
} bkma_dtype;

/* ADVParams, src/config/AdvectionParams.hpp:10-57 (POD device copy).
 * Derived fields (dx, dvx, inv_dx, realWidthX) follow update_deltas,
 * src/config/AdvectionParams.cpp:96-103; see bkma_adv_update_deltas. */
typedef struct bkma_adv_params {
    uint64_t n0, n1, n2;
    uint64_t maxIter;
    uint64_t pref_wg_size;
    uint64_t seq_size0, seq_size2;
    double dt, dx, dvx, inv_dx;
    double minRealX, maxRealX, realWidthX;
    double minRealVx, maxRealVx;
    float percent_loc;
    int32_t gpu;
    int32_t outputSolution;
    int32_t _pad;
} bkma_adv_params;

/* BatchConfig1D, src/core/bkma_tools.hpp:113-122 */
typedef struct bkma_batch_config {
    uint64_t n_batch, batch_size, last_batch_size;
} bkma_batch_config;

/* WorkGroupDispatch, src/core/bkma_tools.hpp:25-56 */
typedef struct bkma_wg_dispatch {
    uint64_t s0, s2; /* sequential sizes */
    uint64_t g0, g2; /* work-groups per batch in dims 0 and 2 */
} bkma_wg_dispatch;

/* BkmaOptimParams, src/core/bkma_tools.hpp:126-134 */
typedef struct bkma_optim_params {
    bkma_batch_config dispatch_d0;
    bkma_batch_config dispatch_d2;
    uint64_t w0, w1, w2;
    bkma_wg_dispatch wg_dispatch;
    int32_t mem_space; /* bkma_memspace */
    int32_t _pad;
} bkma_optim_params;

/* A caller-provided global scratch, the span3d_t `global_scratch` argument of
 * bkma_run (src/core/bkma_run.hpp:10-13).  data == NULL means "empty span". */
typedef struct bkma_span3d {
    void *data;
    uint64_t e0, e1, e2;
} bkma_span3d;

/* ------------------------------------------------------------------ */
/* Library / device                                                     */
/* ------------------------------------------------------------------ */
int bkma_abi_version(void);
const char *bkma_last_error(void);
/* Maximum LDS bytes one work-group may allocate on `device` (the value the
 * reference reads from sycl::info::device::local_mem_size, init.hpp:84-86). */
bkma_status bkma_query_lds_bytes(int device, uint64_t *bytes);

/* ------------------------------------------------------------------ */
/* Parameters and dispatch policy (host arithmetic, no device work)     */
/* ------------------------------------------------------------------ */
/* ADVParams::update_deltas, src/config/AdvectionParams.cpp:96-103 */
void bkma_adv_update_deltas(bkma_adv_params *p);

/* create_optim_params<Params>, src/tools/init.hpp:76-102, composed of
 * WorkItemDispatch::set_ideal_sizes / adjust_sizes_mem_limit
 * (src/core/bkma_tools.hpp:74-107) and WorkGroupDispatch::set_num_work_groups
 * (src/core/bkma_tools.hpp:34-55).  `local_mem_bytes` = 0 queries the current
 * device.  `elem_size` is sizeof the scratch element (8 for the reference's
 * real_t = double).  Returns BKMA_ERR_INVALID_ARGUMENT with the reference's
 * messages where set_num_work_groups throws, and where the reference divides
 * by zero (a line longer than the LDS, bkma_tools.hpp:102-105) it falls back
 * to MemorySpace::Global with w0=w2=1 instead. */
bkma_status bkma_create_optim_params(uint64_t n0, uint64_t n1, uint64_t n2, uint64_t pref_wg_size,
                                     uint64_t seq_size0, uint64_t seq_size2, uint64_t elem_size,
                                     uint64_t local_mem_bytes, bkma_optim_params *out);

/* init_1d_blocking, src/core/bkma_tools.hpp:138-153 */
bkma_batch_config bkma_init_1d_blocking(uint64_t n, uint64_t max_batchs);

/* Elements of global scratch that bkma_adv_run / bkma_conv_run need for
 * (impl, optim) on an [n0][n1][n2] array with the given solver window.
 * AdaptiveWg/Global: extents (g0*w0, g2*w2, n1-window+1), indexed by global
 * id (src/core/MemorySpace.hpp:45-70).  BasicRange: (n0, n1, n2). */
bkma_status bkma_scratch_extents(bkma_impl impl, const bkma_optim_params *optim, uint64_t n0, uint64_t n1,
                                 uint64_t n2, uint64_t window, uint64_t *e0, uint64_t *e1, uint64_t *e2);

/* ------------------------------------------------------------------ */
/* bkma_run<AdvectionSolver, Impl>  (src/core/bkma_run.hpp:10-62)      */
/* ------------------------------------------------------------------ */
/* One in-place advection step of `data` ([n0][n1][n2], dtype F64 or F32).
 * The solver is AdvectionSolver (src/solvers/AdvectionSolver.hpp:13-99):
 * FP64 foot location, lag_basis and stencil sum in the reference's exact
 * operation order; F32 storage is converted on load and rounded on store.
 * impl/optim/scratch have the reference's meaning: AdaptiveWg uses
 * optim->mem_space (Local: LDS scratch; Global: `scratch`); NDRange needs
 * n1 <= 1024; BasicRange needs a (n0, n1, n2) global scratch. */
bkma_status bkma_adv_run(bkma_impl impl, bkma_dtype dtype, void *data, const bkma_adv_params *params,
                         const bkma_optim_params *optim, bkma_span3d scratch, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): bkma_adv_run with the global scratch of AdaptiveWg/Global and
 * BasicRange taken from a library-owned device array of
 * bkma_static_scratch_bytes() bytes (64 MiB) per device, held until the
 * process ends, instead of a caller span.  Calls no hipMalloc/hipFree.
 * A shape whose scratch exceeds the array is rejected with
 * BKMA_ERR_INVALID_ARGUMENT.  Every call on one device shares that device's
 * array: calls that overlap on one GPU (several streams or host threads)
 * race.  Local/NDRange calls need no scratch and behave as bkma_adv_run. */
bkma_status bkma_adv_run_noalloc(bkma_impl impl, bkma_dtype dtype, void *data, const bkma_adv_params *params,
                                 const bkma_optim_params *optim, void *stream);
uint64_t bkma_static_scratch_bytes(void);

/* ------------------------------------------------------------------ */
/* bkma_run<ConvSolver, Impl>  (src/core/bkma_run.hpp:10-62)           */
/* ------------------------------------------------------------------ */
/* ConvSolver (src/solvers/ConvSolver.hpp:7-45) through the BKMA executor,
 * window() = K: data is [n0][n1 = C*L][n2]; weight is the (K, C_in, C_out)
 * span of conv1d.cpp:46 with C_in = C_out = C; bias is [C].  Same dtype for
 * data, weight and bias (F64 or F32).  Writes that the reference performs at
 * negative window indices (outside the line's scratch) are not performed. */
bkma_status bkma_conv_run(bkma_impl impl, bkma_dtype dtype, void *data, uint64_t n0, uint64_t n1, uint64_t n2,
                          const void *weight, const void *bias, uint64_t K, uint64_t C, uint64_t L,
                          const bkma_optim_params *optim, bkma_span3d scratch, void *stream);

/* ------------------------------------------------------------------ */
/* Tuned advection plans                                               */
/* ------------------------------------------------------------------ */
/* Replaces the per-step `bkma_run_function(Q, data, solver, optim_params,
 * span3d_t{})` call of the advection time loop (src/advection.cpp:51-58)
 * for callers that want the tuned kernels: a plan is built once per
 * (ADVParams, storage dtype), chooses a strategy and its launch shape
 * (hand-set in the reference through BkmaOptimParams / AdaptiveWg), then
 * runs any number of in-place steps.
 *  - F64 storage: results bit-identical to bkma_adv_run (the reference's
 *    FP64 arithmetic, reordered only across cells), whatever the number of
 *    steps fused per pass.
 *  - F32 storage: FP64 foot location / index / lag_basis exactly as the
 *    reference, FP64 fused accumulation; the steps fused into one pass keep
 *    their intermediate values in FP64 and round to FP32 once per pass.  A
 *    run of nsteps makes nsteps / steps_per_pass full passes, then one pass
 *    of the remaining steps.  The result therefore depends on how a caller
 *    splits its steps into calls: one call per step (the reference's time
 *    loop) rounds every step, which on near-identity lines loses updates
 *    below half an FP32 ulp each time (rel_max 1.8e-6 against FP64 after 55
 *    steps of [2048][1024][512], bound nsteps * 2^-24); the whole run in one
 *    call stays within 1e-6 (3.5e-7 there with 5-step passes). */
typedef enum bkma_strategy {
    BKMA_STRATEGY_COLUMN_STREAM = 1, /* lines strided by n2 (n2 >= 64), small shifts: register ring per wavefront */
    BKMA_STRATEGY_LINE_LDS = 2,      /* any n2 and shift: slab segments + halos staged in LDS */
    BKMA_STRATEGY_COLUMN_PAIR = 3,   /* as COLUMN_STREAM, for slabs whose rows take at most two adjacent shifts
                                        (checked per slab at plan creation): seven-row windows, no shift branches */
    BKMA_STRATEGY_PLAIN = 4,         /* the reference's AdaptiveWg executor and create_optim_params dispatch
                                        (F64 storage, line in LDS; seg_rows = pref_wg_size), one step per pass */
    BKMA_STRATEGY_HYBRID = 5,        /* the reference's local/global split (HybridMem.cpp:4-32, percent_loc of
                                        AdvectionParams.hpp:18): the first floor(n0 * chunk_segs / 100) slabs run
                                        the AdaptiveWg executor with the line scratch in LDS, the other slabs with
                                        a global scratch, concurrently on a second stream (F64 storage;
                                        seg_rows = pref_wg_size), one step per pass */
    BKMA_STRATEGY_LINE_SHFL = 6      /* contiguous lines (n2 = 1) held in registers for up to 256 steps per pass:
                                        lines of 64 vec k cells (k = 1 .. 16, at most 4096 cells) whose rows take
                                        at most two adjacent shifts, one work-group per line (vec = 4 or 8 cells
                                        per lane, 0 = 8; waves_per_wg = k, at most 8 with 8 cells), neighbours
                                        through DPP wavefront shifts, shifts up to about vec - 3 cells; or lines
                                        of 16 cells (vec = 1), one per 16-lane row, neighbours through DPP row
                                        rotations, all lines' first-tap offsets within 3 consecutive values */
} bkma_strategy;

typedef struct bkma_tuned_config {
    int32_t strategy;     /* bkma_strategy */
    int32_t vec;          /* consecutive i2 cells per lane (LINE_SHFL: consecutive i1 cells per lane) */
    int32_t smax;         /* largest |v*dt/dx| the stencil window covers */
    int32_t prefetch;     /* rows loaded ahead of the stencil window */
    int32_t waves_per_wg; /* wavefronts per work-group (LINE_LDS: 4, 8 or 16; 0 = 4) */
    int32_t nontemporal;  /* 1: non-temporal global loads/stores */
    int32_t persistent;   /* 1: grid capped at the resident work-groups, waves loop over units */
    int32_t seg_rows;     /* LINE_LDS: rows per segment (0 = whole line in LDS); PLAIN, HYBRID: pref_wg_size */
    int32_t wide_window;  /* COLUMN_STREAM, F32 storage: 1 = FP64 register window (rows widened once) */
    int32_t steps_per_pass; /* COLUMN_STREAM, COLUMN_PAIR, LINE_LDS (at most 4), LINE_SHFL (at most 256): time
                               steps fused into one pass over HBM (0 or 1: one) */
    int32_t chunk_segs;   /* LINE_LDS segmented: consecutive segments one work-group walks in order (0 or 1: one);
                             LINE_LDS whole line: consecutive slabs one work-group stages together;
                             HYBRID: percent of the slabs whose line scratch is in LDS (0..100) */
    int32_t dpp_broadcast; /* COLUMN_PAIR: 1 = the coefficients reach the stencil sums through the DPP row
                              broadcast; 0 = through v_readlane */
} bkma_tuned_config;

typedef struct bkma_adv_plan bkma_adv_plan;

/* autotune = 0: heuristic choice; 1: time the feasible variants on a scratch
 * array of the same line shape (allocated and freed inside the call). */
bkma_status bkma_adv_plan_create(const bkma_adv_params *params, bkma_dtype dtype, int32_t autotune, void *stream,
                                 bkma_adv_plan **out);
/* A plan with an explicit configuration (BKMA_ERR_UNSUPPORTED if it is not
 * compiled in or cannot cover the shape). */
bkma_status bkma_adv_plan_create_with(const bkma_adv_params *params, bkma_dtype dtype,
                                      const bkma_tuned_config *cfg, bkma_adv_plan **out);
bkma_status bkma_adv_plan_describe(const bkma_adv_plan *plan, bkma_tuned_config *cfg, char *name,
                                   uint64_t name_len);
/* nsteps in-place steps on `data` (16-byte aligned, the plan's dtype). */
bkma_status bkma_adv_plan_run(bkma_adv_plan *plan, void *data, uint64_t nsteps, void *stream);
void bkma_adv_plan_destroy(bkma_adv_plan *plan);
/* The compiled configurations of `dtype` (the autotuner's search space):
 * writes up to max_out of them and returns how many exist. */
int32_t bkma_adv_plan_variants(bkma_dtype dtype, bkma_tuned_config *out, int32_t max_out);

/* ------------------------------------------------------------------ */
/* Ragged batches of advection problems                                */
/* ------------------------------------------------------------------ */
/* Many independent problems (each its own [n0][n1][n2] array, dt and
 * velocity grid; the reference runs one bkma_run per problem and step)
 * advanced together: every pass of ns steps (the batch's steps per pass,
 * 1 .. 4) is one persistent launch over all problems' work items (plus a
 * halo-snapshot launch when a problem's lines exceed the LDS budget), and
 * bkma_adv_batch_run(use_graph = 1) captures the passes into hipGraphs,
 * replayed on later calls.  A run of nsteps steps makes nsteps / ns passes of
 * ns steps, then one pass of the rest.  Results are bit-identical to running
 * each problem's passes alone (FP64: the reference's arithmetic, whatever
 * ns; FP32: the tuned FP32 arithmetic with one rounding to FP32 per pass). */
typedef struct bkma_adv_problem {
    void *data;              /* device array [n0][n1][n2] of the batch dtype, 16-byte aligned */
    bkma_adv_params params;  /* update_deltas applied */
} bkma_adv_problem;

typedef struct bkma_adv_batch bkma_adv_batch;

/* autotune = 1 times the LDS budgets (32 KB .. the device limit), work-group
 * sizes and steps per pass (1 .. 4) on the problems' own arrays, which the
 * timing steps modify: create the batch before initialising the data.
 * autotune = 0: 64 KB, 256 threads, one step per pass. */
bkma_status bkma_adv_batch_create(const bkma_adv_problem *problems, int32_t nproblems, bkma_dtype dtype,
                                  int32_t autotune, void *stream, bkma_adv_batch **out);
/* A batch with an explicit configuration: LDS budget per work-group (bytes),
 * threads per work-group (256 or 1024; steps_per_pass > 1 needs 256) and
 * steps per pass (1 .. 4).  BKMA_ERR_UNSUPPORTED when a problem does not fit. */
bkma_status bkma_adv_batch_create_with(const bkma_adv_problem *problems, int32_t nproblems, bkma_dtype dtype,
                                       uint64_t lds_budget, int32_t threads, int32_t steps_per_pass, void *stream,
                                       bkma_adv_batch **out);
bkma_status bkma_adv_batch_run(bkma_adv_batch *batch, uint64_t nsteps, int32_t use_graph, void *stream);
/* Baselines over the same problems: mode 0 = the reference's loop of
 * per-problem bkma_run<AdvectionSolver, AdaptiveWg> launches with
 * create_optim_params (pref_wg_size 256); mode 1 = per-problem tuned plans. */
bkma_status bkma_adv_batch_run_loop(bkma_adv_batch *batch, uint64_t nsteps, int32_t mode, void *stream);
bkma_status bkma_adv_batch_describe(const bkma_adv_batch *batch, uint64_t *cells, uint64_t *items,
                                    uint64_t *segmented_items, uint64_t *lds_budget, uint32_t *grid);
/* Steps per pass of the batch (0 for a null batch). */
int32_t bkma_adv_batch_steps_per_pass(const bkma_adv_batch *batch);
void bkma_adv_batch_destroy(bkma_adv_batch *batch);

/* ------------------------------------------------------------------ */
/* Conv1d, textbook semantics (torch.nn.functional.conv1d, stride 1,   */
/* no padding), in place                                                */
/* ------------------------------------------------------------------ */
/* x[B][C][L] (contiguous, the reference's [n0][n1 = C*L][n2 = 1] BKMA
 * layout), weight[C_out = C][C_in = C][K], bias[C].  One run writes
 * y[b][oc][j], j < L-K+1, over x[b][oc][j] and keeps x[b][oc][L-K+1 ..).
 * Replaces the ConvSolver call of the reference's conv1d driver
 * (src/conv1d.cpp:83-87) for callers that want the textbook contraction
 * (the reference functor's own semantics stay available as bkma_conv_run).
 *  BKMA_CONV_F64   : io F64, weights F64; FP64, the oracle's summation order
 *  BKMA_CONV_F32   : io F32, weights F32; FP32 fused multiply-add
 *  BKMA_CONV_BF16  : io F32 or BF16, weights F32; bf16 MFMA, FP32 accumulate
 *  BKMA_CONV_BF16X3: io F32 or BF16, weights F32; split bf16 (hi*hi + hi*lo
 *                    + lo*hi) MFMA, FP32 accumulate
 *  BKMA_CONV_BF16X6: io F32, weights F32; three-term bf16 split (x = x0+x1+x2,
 *                    w = w0+w1+w2 exactly), the six products down to 2^-16
 *                    on bf16 MFMA, FP32 accumulation: FP32-level accuracy
 *  BKMA_CONV_FP8   : io F32 or BF16, weights F32; OCP FP8 e4m3 MFMA, FP32
 *                    accumulate.  Weights are quantized per output channel
 *                    (scale = the power of two >= max|w[oc]| / 448); the
 *                    input is quantized as x / input_scale (per tensor, set
 *                    with bkma_conv1d_plan_set_input_scale, default 1). */
typedef enum bkma_conv_mode {
    BKMA_CONV_F64 = 0,
    BKMA_CONV_F32 = 1,
    BKMA_CONV_BF16 = 2,
    BKMA_CONV_BF16X3 = 3,
    BKMA_CONV_FP8 = 4,
    BKMA_CONV_BF16X6 = 5
} bkma_conv_mode;

typedef struct bkma_conv1d_plan bkma_conv1d_plan;

bkma_status bkma_conv1d_plan_create(bkma_conv_mode mode, bkma_dtype io_dtype, uint64_t B, uint64_t C, uint64_t L,
                                    uint64_t K, bkma_dtype w_dtype, const void *weight, const void *bias,
                                    void *stream, bkma_conv1d_plan **out);
bkma_status bkma_conv1d_plan_run(bkma_conv1d_plan *plan, void *x, void *stream);
bkma_status bkma_conv1d_plan_describe(const bkma_conv1d_plan *plan, char *name, uint64_t name_len);
/* FP8 mode: the input is quantized as x / scale (scale > 0; default 1). */
bkma_status bkma_conv1d_plan_set_input_scale(bkma_conv1d_plan *plan, float scale);
void bkma_conv1d_plan_destroy(bkma_conv1d_plan *plan);

/* ------------------------------------------------------------------ */
/* Harness kernels: fill_buffer_* (src/tools/init.hpp:31-72) and the   */
/* validators (src/tools/validation.hpp)                                */
/* ------------------------------------------------------------------ */
/* U[0,1) from the counter-based generator keyed by (seed, linear index);
 * bit-identical to the oracle's generator (53 random bits, then rounded to
 * the storage type). */
bkma_status bkma_fill_uniform(bkma_dtype dtype, void *out, uint64_t n, uint64_t seed, void *stream);
/* N(0, stddev^2) by Box-Muller over two generator draws per element. */
bkma_status bkma_fill_normal(bkma_dtype dtype, void *out, uint64_t n, uint64_t seed, double stddev, void *stream);
/* Every element set to `value` (rounded to the storage type); the constant
 * fills of fill_buffer_conv1d, src/tools/init.hpp:50-72. */
bkma_status bkma_fill_constant(bkma_dtype dtype, void *out, uint64_t n, double value, void *stream);
/* fill_buffer_adv, src/tools/init.hpp:31-46: sin(4*x*pi). */
bkma_status bkma_fill_adv_sin4(bkma_dtype dtype, void *data, const bkma_adv_params *params, void *stream);
/* validate_result_adv, src/tools/validation.hpp:9-80 (synchronous):
 * highest / lowest over i2 of mean_(i0,i1) |f - sin(4 pi (x - v t))|,
 * t = params->maxIter * dt. */
bkma_status bkma_validate_adv_l1(bkma_dtype dtype, const void *data, const bkma_adv_params *params,
                                 double *highest, double *lowest, void *stream);
/* sum_and_normalize_conv1d, src/tools/validation.hpp:84-110 (synchronous). */
bkma_status bkma_conv_sum_normalize(bkma_dtype dtype, const void *data, uint64_t n0, uint64_t n1, uint64_t n2,
                                    uint64_t nw, double *mean, void *stream);
/* validate_conv1d, src/tools/validation.hpp:114-156 (synchronous):
 * *all_equal = 1 when data(:, i1, :) is constant for every i1 < nw. */
bkma_status bkma_conv_validate(bkma_dtype dtype, const void *data, uint64_t n0, uint64_t n1, uint64_t n2,
                               uint64_t nw, int32_t *all_equal, void *stream);
/* Error norms of `a` against `ref` over n elements (synchronous):
 * out[0] = max|a-ref|, out[1] = max|ref|, out[2] = sum (a-ref)^2,
 * out[3] = sum ref^2, all accumulated in FP64. */
bkma_status bkma_error_norms(bkma_dtype a_dtype, const void *a, bkma_dtype ref_dtype, const void *ref, uint64_t n,
                             double out[4], void *stream);

#ifdef __cplusplus
}
#endif

#endif /* BKMA_AMD_H */
