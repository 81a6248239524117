// Harness helpers of the reference's src/tools/init.hpp and validation.hpp:
// device selection, buffer fills, validators and PERF_DIAGS printing.
#pragma once

#include <chrono>
#include <cmath>
#include <iostream>

#include "bkma.hpp"

namespace bkma {

// pick_device (src/tools/init.hpp:9-27).  There is no CPU device here: asking
// for the CPU, or for a GPU on a host without one, is an error (the
// reference falls back to its SYCL CPU device).
inline int pick_device(bool run_on_gpu) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        throw device_error("no HIP GPU available; this library has no CPU device");
    if (!run_on_gpu)
        std::cout << "CPU execution was requested; the kernels run on GPU 0\n" << std::endl;
    return 0;
}

// fill_buffer_adv (src/tools/init.hpp:31-46): sin(4 * x * pi).
template <class T> void fill_buffer_adv(Queue &q, span3d<T> &data, const ADVParams &params) {
    ADVParams p = params;
    p.n0 = data.extent(0);
    p.n1 = data.extent(1);
    p.n2 = data.extent(2);
    throw_on(bkma_fill_adv_sin4(dtype_of<T>(), data.ptr, &p, q.native()));
    q.wait();
}

// fill_buffer_conv1d (src/tools/init.hpp:50-72): data 7.3, warm-up data 1.0,
// weights 1.5, bias 1.0.
template <class T>
void fill_buffer_conv1d(Queue &q, span3d<T> &data, span3d<T> &warmup_data, span3d<T> &weight, span1d<T> &bias) {
    throw_on(bkma_fill_constant(dtype_of<T>(), data.ptr, data.size(), 7.3, q.native()));
    throw_on(bkma_fill_constant(dtype_of<T>(), warmup_data.ptr, warmup_data.size(), 1.0, q.native()));
    throw_on(bkma_fill_constant(dtype_of<T>(), weight.ptr, weight.size(), 1.5, q.native()));
    throw_on(bkma_fill_constant(dtype_of<T>(), bias.ptr, bias.size(), 1.0, q.native()));
    q.wait();
}

// validate_result_adv (src/tools/validation.hpp:9-80).
template <class T>
double validate_result_adv(Queue &q, span3d<T> &data, const ADVParams &params, bool do_print = true) {
    if (do_print) std::cout << "\nRESULTS_VALIDATION:" << std::endl;
    ADVParams p = params;
    p.n0 = data.extent(0);
    p.n1 = data.extent(1);
    p.n2 = data.extent(2);
    double hi = 0, lo = 0;
    throw_on(bkma_validate_adv_l1(dtype_of<T>(), data.ptr, &p, &hi, &lo, q.native()));
    if (do_print) std::cout << "Highest L1 error found: " << hi << " (lowest is " << lo << ")\n" << std::endl;
    return hi;
}

// sum_and_normalize_conv1d (src/tools/validation.hpp:84-110).
template <class T> double sum_and_normalize_conv1d(Queue &q, span3d<T> data, size_t nw) {
    double mean = 0;
    throw_on(bkma_conv_sum_normalize(dtype_of<T>(), data.ptr, data.extent(0), data.extent(1), data.extent(2), nw,
                                     &mean, q.native()));
    return mean;
}

// validate_conv1d (src/tools/validation.hpp:114-156).
template <class T> void validate_conv1d(Queue &q, span3d<T> &data, size_t nw) {
    int32_t eq = 0;
    throw_on(bkma_conv_validate(dtype_of<T>(), data.ptr, data.extent(0), data.extent(1), data.extent(2), nw, &eq,
                                q.native()));
    if (!eq)
        std::cout << "WARNING: Values at same position i1 are not equivalent throught the batchs. Check "
                     "implementation."
                  << std::endl;
    else
        std::cout << "All values data[:,i1,:] are equal." << std::endl;
}

// print_perf (src/tools/validation.hpp:160-169), with the element size of
// the storage type in the GB/s estimate (2 * sizeof(T) per cell update).
inline void print_perf(double elapsed_seconds, size_t n_cells, size_t elem_size = sizeof(real_t)) {
    std::cout << "PERF_DIAGS:" << std::endl;
    std::cout << "elapsed_time: " << elapsed_seconds << " s\n";
    const double gcells = ((double)n_cells / elapsed_seconds) / 1e9;
    std::cout << "upd_cells_per_sec: " << gcells << " Gcell/sec\n";
    std::cout << "estimated_throughput: " << gcells * (double)elem_size * 2 << " GB/s" << std::endl;
}

}  // namespace bkma
