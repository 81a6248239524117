// Two solvers that are not the library's, run through the C++ facade's
// bkma_run / impl_selector and the header-only executors (bkma/plugin.hpp):
//  - ShiftSolver (window 1): row i1 takes row (i1 - s(i0)) mod n1, s(i0) = i0 % 5 + 1,
//    so a step is a circular shift of every (i0, :, i2) line by s(i0);
//  - WindowSumSolver (window W): row i1 >= W-1 gives the sum of rows i1-W+1 .. i1,
//    written back at i1-(W-1); rows nw = n1-W+1 .. n1-1 keep their values.
// TEST INFRASTRUCTURE: built into tests/cpp/libplugin_solvers.so by
// tests/cpp/Makefile and called by tests/test_plugin_gpu.py.
#include <cstring>
#include <string>

#include "bkma/plugin.hpp"

namespace {

using i64 = long long;

struct ShiftSolver {
    i64 n1;
    __host__ __device__ i64 window() const { return 1; }
    template <class Line> __device__ double operator()(const Line &line, i64 i0, i64 i1, i64) const {
        const i64 s = i0 % 5 + 1;
        return line((i1 + n1 - s) % n1);
    }
};

struct WindowSumSolver {
    i64 w;
    __host__ __device__ i64 window() const { return w; }
    template <class Line> __device__ double operator()(const Line &line, i64, i64 i1, i64) const {
        double acc = 0.;
        for (i64 k = 0; k < w; ++k) acc += line(i1 - k);
        return acc;
    }
};

void put_error(const std::exception &e, char *err, int len) {
    std::strncpy(err, e.what(), (size_t)len - 1);
    err[len - 1] = '\0';
}

template <class T, class Solver>
void run_one(int impl, T *ptr, size_t n0, size_t n1, size_t n2, const Solver &solver, int global_scratch) {
    bkma::Queue Q(0);
    bkma::span3d<T> data(ptr, n0, n1, n2);
    bkma::ADVParams shape;  // create_optim_params reads n0, n1, n2, pref_wg_size and the seq sizes
    shape.n0 = n0;
    shape.n1 = n1;
    shape.n2 = n2;
    shape.pref_wg_size = 128;
    auto optim = bkma::create_optim_params(Q, shape, sizeof(T));
    bkma::span3d<T> scratch;
    T *owned = nullptr;
    if (impl == 2) {  // BasicRange: a (n0, n1, n2) scratch
        owned = bkma::device_alloc<T>(n0 * n1 * n2);
        scratch = bkma::span3d<T>(owned, n0, n1, n2);
        bkma::bkma_run<Solver, bkma::BkmaImpl::BasicRange, T>(Q, data, solver, optim, scratch).wait();
    } else if (impl == 1) {  // NDRange, selected by name
        auto fn = bkma::impl_selector<Solver, T>("NDRange");
        fn(Q, data, solver, optim, scratch).wait();
    } else {  // AdaptiveWg, selected by name, LDS or global scratch
        if (global_scratch) {
            optim.mem_space = BKMA_MEM_GLOBAL;
            uint64_t e0, e1, e2;
            bkma::throw_on(bkma_scratch_extents(BKMA_IMPL_ADAPTIVE_WG, &optim, n0, n1, n2, (uint64_t)solver.window(),
                                                &e0, &e1, &e2));
            owned = bkma::device_alloc<T>(e0 * e1 * e2);
            scratch = bkma::span3d<T>(owned, e0, e1, e2);
        }
        auto fn = bkma::impl_selector<Solver, T>("adaptiveWG");
        fn(Q, data, solver, optim, scratch).wait();
    }
    bkma::device_free(owned);
}

}  // namespace

extern "C" {
// solver 0 = ShiftSolver, 1 = WindowSumSolver(window); impl 0 = AdaptiveWg,
// 1 = NDRange, 2 = BasicRange; f64 = 1 for double data, 0 for float.
// Returns 0, or 1 with the exception's message in err.
int plugin_run(int solver, int impl, int f64, void *data, unsigned long long n0, unsigned long long n1,
               unsigned long long n2, long long window, int global_scratch, char *err, int errlen) {
    try {
        if (solver == 0) {
            const ShiftSolver s{(i64)n1};
            if (f64)
                run_one(impl, (double *)data, n0, n1, n2, s, global_scratch);
            else
                run_one(impl, (float *)data, n0, n1, n2, s, global_scratch);
        } else {
            const WindowSumSolver s{window};
            if (f64)
                run_one(impl, (double *)data, n0, n1, n2, s, global_scratch);
            else
                run_one(impl, (float *)data, n0, n1, n2, s, global_scratch);
        }
        return 0;
    } catch (const std::exception &e) {
        put_error(e, err, errlen);
        return 1;
    }
}
}
