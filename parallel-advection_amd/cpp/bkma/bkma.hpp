// C++ host interface of the gfx950 BKMA library, shaped like the reference's
// src/core + src/solvers + src/tools interface so that a caller of
// bkma_run<Solver, Impl>, impl_selector, create_optim_params and the solver
// functors keeps its code:
//
//   reference (SYCL)                       here (HIP, over include/bkma_amd.h)
//   sycl::queue                            bkma::Queue   (a HIP stream)
//   sycl::event                            bkma::Event   (a HIP event)
//   span3d_t / span1d_t (mdspan, real_t)   bkma::span3d_t / span1d_t (device view, any T)
//   bkma_run<Solver, Impl>(Q, data, solver, optim, scratch)   same
//   impl_selector<Solver>(name)            same (lower-cased names, same error text)
//   create_optim_params<Params>(Q, params) same policy (bkma_create_optim_params)
//   AdvectionSolver{params}, ConvSolver{w, b, k, c_in, L}     same members
//
// Errors from the C ABI are rethrown with the reference's exception types:
// std::invalid_argument, std::runtime_error, and bkma::device_error for HIP
// failures (where the reference would throw sycl::exception).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "bkma_amd.h"
#include "config.hpp"

namespace bkma {

struct device_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// A failing status as the reference's exception type.
[[noreturn]] inline void throw_status(bkma_status st, const std::string &msg) {
    switch (st) {
        case BKMA_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case BKMA_ERR_DEVICE: throw device_error(msg);
        default: throw std::runtime_error(msg);
    }
}

inline void throw_on(bkma_status st) {
    if (st != BKMA_OK) throw_status(st, bkma_last_error());
}

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw device_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Queue / Event
// ---------------------------------------------------------------------------
class Event {
  public:
    Event() = default;
    explicit Event(hipStream_t s) {
        hip_check(hipEventCreateWithFlags(&ev_, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventRecord(ev_, s), "hipEventRecord");
    }
    Event(const Event &) = delete;
    Event &operator=(const Event &) = delete;
    Event(Event &&o) noexcept : ev_(o.ev_) { o.ev_ = nullptr; }
    Event &operator=(Event &&o) noexcept {
        if (this != &o) {
            reset();
            ev_ = o.ev_;
            o.ev_ = nullptr;
        }
        return *this;
    }
    ~Event() { reset(); }
    void wait() const {
        if (ev_) hip_check(hipEventSynchronize(ev_), "hipEventSynchronize");
    }

  private:
    void reset() {
        if (ev_) (void)hipEventDestroy(ev_);
        ev_ = nullptr;
    }
    hipEvent_t ev_ = nullptr;
};

// An in-order queue on one device (the reference's queues are waited on
// after each submission, so in-order execution preserves their semantics).
class Queue {
  public:
    explicit Queue(int device = 0) : dev_(device) {
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate");
    }
    Queue(const Queue &) = delete;
    Queue &operator=(const Queue &) = delete;
    ~Queue() {
        if (s_) (void)hipStreamDestroy(s_);
    }
    void wait() const { hip_check(hipStreamSynchronize(s_), "hipStreamSynchronize"); }
    hipStream_t stream() const { return s_; }
    void *native() const { return reinterpret_cast<void *>(s_); }
    int device() const { return dev_; }
    std::string device_name() const {
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev_), "hipGetDeviceProperties");
        return prop.name + std::string(" (") + prop.gcnArchName + ")";
    }
    size_t local_mem_size() const {
        uint64_t b = 0;
        throw_on(bkma_query_lds_bytes(dev_, &b));
        return (size_t)b;
    }

  private:
    int dev_ = 0;
    hipStream_t s_ = nullptr;
};

// ---------------------------------------------------------------------------
// Device views (row-major, layout_right) and allocation
// ---------------------------------------------------------------------------
template <class T> struct span1d {
    T *ptr = nullptr;
    size_t n = 0;
    size_t extent(int) const { return n; }
    T *data_handle() const { return ptr; }
    size_t size() const { return n; }
};

template <class T> struct span3d {
    T *ptr = nullptr;
    size_t e[3] = {0, 0, 0};
    span3d() = default;
    span3d(T *p, size_t e0, size_t e1, size_t e2) : ptr(p), e{e0, e1, e2} {}
    size_t extent(int i) const { return e[i]; }
    T *data_handle() const { return ptr; }
    size_t size() const { return e[0] * e[1] * e[2]; }
    bool empty() const { return ptr == nullptr; }
};

using real_t = double;  // the reference's compute type (src/types.hpp:5)
using span1d_t = span1d<real_t>;
using span3d_t = span3d<real_t>;

template <class T> constexpr bkma_dtype dtype_of();
template <> constexpr bkma_dtype dtype_of<double>() { return BKMA_F64; }
template <> constexpr bkma_dtype dtype_of<float>() { return BKMA_F32; }

// sycl_alloc (src/types.hpp:7-10): device memory for n elements.
template <class T = real_t> T *device_alloc(size_t n) {
    void *p = nullptr;
    hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    return static_cast<T *>(p);
}
inline void device_free(void *p) {
    if (p) (void)hipFree(p);
}
inline real_t *sycl_alloc(size_t n, Queue &) { return device_alloc<real_t>(n); }

// ---------------------------------------------------------------------------
// Dispatch types (src/core/bkma_tools.hpp, MemorySpace.hpp)
// ---------------------------------------------------------------------------
enum class BkmaImpl { BasicRange, NDRange, AdaptiveWg };
enum class MemorySpace { Local, Global };

using BatchConfig1D = bkma_batch_config;
using BkmaOptimParams = bkma_optim_params;

inline BatchConfig1D init_1d_blocking(size_t n, size_t max_batchs) { return bkma_init_1d_blocking(n, max_batchs); }

inline MemorySpace mem_space_of(const BkmaOptimParams &o) {
    return o.mem_space == BKMA_MEM_GLOBAL ? MemorySpace::Global : MemorySpace::Local;
}

// create_optim_params<Params> (src/tools/init.hpp:76-102); elem_size is the
// scratch element size (sizeof(real_t) in the reference).
template <class Params>
BkmaOptimParams create_optim_params(Queue &q, const Params &params, size_t elem_size = sizeof(real_t)) {
    BkmaOptimParams o;
    throw_on(bkma_create_optim_params(params.n0, params.n1, params.n2, params.pref_wg_size, params.seq_size0,
                                      params.seq_size2, elem_size, q.local_mem_size(), &o));
    return o;
}

// ---------------------------------------------------------------------------
// Solvers (src/solvers/)
// ---------------------------------------------------------------------------
struct AdvectionSolver {
    ADVParams params;
    AdvectionSolver() = delete;
    explicit AdvectionSolver(const ADVParams &p) : params(p) {}
    int window() const { return 1; }
};

template <class T = real_t> struct ConvSolverT {
    span3d<T> weight_span_;  // (K, C_in, C_out)
    span1d<T> bias_span_;    // (C_out)
    size_t kernel_size_;
    size_t in_channels_;
    size_t input_length_;
    static constexpr size_t stride_ = 1;
    static constexpr size_t padding_ = 0;
    size_t window() const { return kernel_size_; }
};
using ConvSolver = ConvSolverT<real_t>;

// ---------------------------------------------------------------------------
// bkma_run<Solver, Impl> (src/core/bkma_run.hpp:10-62)
// ---------------------------------------------------------------------------
constexpr bkma_impl to_c(BkmaImpl i) {
    return i == BkmaImpl::BasicRange ? BKMA_IMPL_BASIC_RANGE
                                     : (i == BkmaImpl::NDRange ? BKMA_IMPL_ND_RANGE : BKMA_IMPL_ADAPTIVE_WG);
}

template <class T> bkma_span3d to_c(const span3d<T> &s) {
    return bkma_span3d{s.ptr, s.e[0], s.e[1], s.e[2]};
}

// The library's two solvers run through the C ABI (their kernels are in
// libbkma_amd.so); any other solver runs through the header-only executors,
// whose generic BkmaRunner is defined in bkma/plugin.hpp (hipcc only).
template <class Solver, BkmaImpl Impl, class T = real_t> struct BkmaRunner;

template <BkmaImpl Impl, class T> struct BkmaRunner<AdvectionSolver, Impl, T> {
    static Event run(Queue &Q, span3d<T> data, const AdvectionSolver &s, BkmaOptimParams optim,
                     span3d<T> scratch) {
        ADVParams p = s.params;
        p.n0 = data.extent(0);
        p.n1 = data.extent(1);
        p.n2 = data.extent(2);
        throw_on(bkma_adv_run(to_c(Impl), dtype_of<T>(), data.ptr, &p, &optim, to_c(scratch), Q.native()));
        return Event(Q.stream());
    }
};

template <BkmaImpl Impl, class T> struct BkmaRunner<ConvSolverT<T>, Impl, T> {
    static Event run(Queue &Q, span3d<T> data, const ConvSolverT<T> &s, BkmaOptimParams optim, span3d<T> scratch) {
        throw_on(bkma_conv_run(to_c(Impl), dtype_of<T>(), data.ptr, data.extent(0), data.extent(1), data.extent(2),
                               s.weight_span_.ptr, s.bias_span_.ptr, s.kernel_size_, s.in_channels_,
                               s.input_length_, &optim, to_c(scratch), Q.native()));
        return Event(Q.stream());
    }
};

template <class Solver, BkmaImpl Impl, class T = real_t>
inline Event bkma_run(Queue &Q, span3d<T> data, const Solver &solver, BkmaOptimParams optim_params,
                      span3d<T> global_scratch = span3d<T>{}) {
    return BkmaRunner<Solver, Impl, T>::run(Q, data, solver, optim_params, global_scratch);
}

// ---------------------------------------------------------------------------
// impl_selector (src/tools/impl_selector.hpp:31-49)
// ---------------------------------------------------------------------------
static constexpr auto error_str = "Should be: {BasicRange, NDRange, AdaptiveWg}";

inline std::string to_lowercase(const std::string &input) {
    std::string r = input;
    std::transform(r.begin(), r.end(), r.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return r;
}

template <class Solver, class T = real_t>
using BkmaRunFn = std::function<Event(Queue &, span3d<T>, const Solver &, BkmaOptimParams, span3d<T>)>;

// Same names, same lower-casing and same error as the reference: NDRange and
// AdaptiveWg are selectable; BasicRange is rejected like any other name (its
// case is commented out at impl_selector.hpp:38-39, since it needs a caller
// scratch) and stays callable as bkma_run<Solver, BkmaImpl::BasicRange>.
template <class Solver, class T = real_t> inline BkmaRunFn<Solver, T> impl_selector(const std::string &impl_name) {
    const std::string impl = to_lowercase(impl_name);
    if (impl == "ndrange") return &bkma_run<Solver, BkmaImpl::NDRange, T>;
    if (impl == "adaptivewg") return &bkma_run<Solver, BkmaImpl::AdaptiveWg, T>;
    throw std::runtime_error(impl_name + " is not a valid implementation name.\n" + error_str);
}

// ---------------------------------------------------------------------------
// Tuned plans (no reference counterpart; see include/bkma_amd.h)
// ---------------------------------------------------------------------------
class TunedAdvection {
  public:
    TunedAdvection(Queue &q, const ADVParams &p, bkma_dtype storage, bool autotune = true) {
        throw_on(bkma_adv_plan_create(&p, storage, autotune ? 1 : 0, q.native(), &plan_));
    }
    TunedAdvection(const TunedAdvection &) = delete;
    TunedAdvection &operator=(const TunedAdvection &) = delete;
    ~TunedAdvection() { bkma_adv_plan_destroy(plan_); }
    template <class T> Event run(Queue &q, span3d<T> data, size_t nsteps = 1) {
        throw_on(bkma_adv_plan_run(plan_, data.ptr, nsteps, q.native()));
        return Event(q.stream());
    }
    std::string name() const {
        char buf[128];
        throw_on(bkma_adv_plan_describe(plan_, nullptr, buf, sizeof(buf)));
        return buf;
    }

  private:
    bkma_adv_plan *plan_ = nullptr;
};

}  // namespace bkma
