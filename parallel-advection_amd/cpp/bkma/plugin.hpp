// bkma_run<Solver, Impl> for a caller's own solver: the generic BkmaRunner
// over the header-only executors of bkma/executors.hpp, so that any
// device functor with the reference's solver contract (window() plus
// operator()(line, i0, i1, i2), src/core/bkma_run.hpp:10-13,
// src/core/AdaptiveWg.hpp:57-64) runs through bkma_run and impl_selector
// like AdvectionSolver and ConvSolver do.  Include it from a translation
// unit compiled with hipcc --offload-arch=gfx950: the executors' kernels are
// instantiated there, for that solver.
//
//   struct Shift {
//       long long n1, s;
//       __host__ __device__ long long window() const { return 1; }
//       template <class Line>
//       __device__ double operator()(const Line &line, long long, long long i1, long long) const {
//           return line((i1 + n1 - s) % n1);
//       }
//   };
//   bkma::bkma_run<Shift, bkma::BkmaImpl::AdaptiveWg>(Q, data, Shift{n1, 3}, optim);
#pragma once

#include "bkma.hpp"
#include "executors.hpp"

namespace bkma {

template <class Solver, BkmaImpl Impl, class T> struct BkmaRunner {
    static Event run(Queue &Q, span3d<T> data, const Solver &solver, BkmaOptimParams optim, span3d<T> scratch) {
        const exec::Status st = exec::run(to_c(Impl), data.ptr, solver, (exec::i64)data.extent(0),
                                          (exec::i64)data.extent(1), (exec::i64)data.extent(2), &optim,
                                          to_c(scratch), Q.stream());
        if (!st.ok()) throw_status(st.code, st.msg);
        return Event(Q.stream());
    }
};

}  // namespace bkma
