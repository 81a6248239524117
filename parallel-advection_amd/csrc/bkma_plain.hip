// The reference's three BKMA executors (BasicRange, NDRange, AdaptiveWg with
// Local or Global scratch) for the library's two solvers, and the
// bkma_adv_run / bkma_conv_run entry points that select them.
//
// These are the "plain" kernels: the header-only executors of
// bkma/executors.hpp (the reference's work decomposition, generic over the
// solver) instantiated with AdvSolverDev and ConvSolverDev, which keep the
// reference's FP64 arithmetic (per-cell foot location and lag_basis), so
// their results are bit-identical to the FP64 oracle.  They are the baseline
// the tuned kernels are measured against.
#include <algorithm>
#include <cstdio>

#include "bkma/executors.hpp"
#include "common.hpp"

namespace bkma {

namespace {

// Status of an executor call as this library's thread-local last error.
bkma_status report(const exec::Status &st) {
    if (st.ok()) return BKMA_OK;
    return fail(st.code, st.msg);
}

template <class T, class Solver>
bkma_status run_impl(bkma_impl impl, T *data, const Solver &solver, i64 n0, i64 n1, i64 n2,
                     const bkma_optim_params *optim, const bkma_span3d &scratch, hipStream_t stream) {
    return report(exec::run(impl, data, solver, n0, n1, n2, optim, scratch, stream));
}

}  // namespace

bkma_status adaptivewg_slabs(double *data, const bkma_adv_params &p, u64 i0_begin, u64 count,
                             const bkma_optim_params &o, const bkma_span3d &scratch, hipStream_t stream) {
    if (count == 0 || i0_begin + count > p.n0) return fail(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg: slab range out of the array");
    if (p.n1 > 0x7fffffffULL || p.n0 > 0x7fffffffULL)
        return fail(BKMA_ERR_UNSUPPORTED, "n0 and n1 must fit in a 32-bit int (the solver's index type)");
    AdvSolverDev solver{make_adv_dev(p)};
    return report(
        exec::run_adaptivewg(data, solver, (i64)count, (i64)p.n1, (i64)p.n2, o, scratch, stream, (i64)i0_begin));
}

}  // namespace bkma

using namespace bkma;

extern "C" bkma_status bkma_adv_run(bkma_impl impl, bkma_dtype dtype, void *data, const bkma_adv_params *params,
                                    const bkma_optim_params *optim, bkma_span3d scratch, void *stream) {
    if (!data || !params) return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_adv_run: null data or params");
    BKMA_TRY(require_device());
    AdvSolverDev solver{make_adv_dev(*params)};
    const i64 n0 = (i64)params->n0, n1 = (i64)params->n1, n2 = (i64)params->n2;
    switch (dtype) {
        case BKMA_F64:
            return run_impl(impl, (double *)data, solver, n0, n1, n2, optim, scratch, as_stream(stream));
        case BKMA_F32:
            return run_impl(impl, (float *)data, solver, n0, n1, n2, optim, scratch, as_stream(stream));
        default:
            return fail(BKMA_ERR_UNSUPPORTED, "bkma_adv_run: dtype must be F64 or F32");
    }
}

extern "C" bkma_status bkma_conv_run(bkma_impl impl, bkma_dtype dtype, void *data, uint64_t n0, uint64_t n1,
                                     uint64_t n2, const void *weight, const void *bias, uint64_t K, uint64_t C,
                                     uint64_t L, const bkma_optim_params *optim, bkma_span3d scratch, void *stream) {
    if (!data || !weight || !bias) return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_conv_run: null pointer");
    if (K == 0 || C == 0 || L == 0) return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_conv_run: K, C, L must be positive");
    if (K > n1) return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_conv_run: kernel size larger than the line");
    // Conv1dParams::setup (src/config/Conv1dParams.cpp:56): n1 = length * channel_out.
    if (n1 != C * L) return fail(BKMA_ERR_INVALID_ARGUMENT, "bkma_conv_run: n1 must equal C * L");
    BKMA_TRY(require_device());
    switch (dtype) {
        case BKMA_F64: {
            ConvSolverDev<double> s{(const double *)weight, (const double *)bias, (i64)K, (i64)C, (i64)L, (i64)n1};
            return run_impl(impl, (double *)data, s, (i64)n0, (i64)n1, (i64)n2, optim, scratch, as_stream(stream));
        }
        case BKMA_F32: {
            ConvSolverDev<float> s{(const float *)weight, (const float *)bias, (i64)K, (i64)C, (i64)L, (i64)n1};
            return run_impl(impl, (float *)data, s, (i64)n0, (i64)n1, (i64)n2, optim, scratch, as_stream(stream));
        }
        default:
            return fail(BKMA_ERR_UNSUPPORTED, "bkma_conv_run: dtype must be F64 or F32");
    }
}
