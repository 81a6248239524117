// conv1d [file.ini]: the reference's Conv1d driver (src/conv1d.cpp:13-111) on
// the gfx950 BKMA library: ConvSolver through bkma_run<ConvSolver, AdaptiveWg>
// with the reference's dispatch policy, constant inputs, warm-up and checks.
#include <chrono>
#include <iostream>

#include "bkma/bkma.hpp"
#include "bkma/tools.hpp"

using namespace bkma;

int main(int argc, char **argv) {
    try {
        const std::string input_file = argc > 1 ? std::string(argv[1]) : "conv1d.ini";
        ConfigMap configMap(input_file);
        Conv1dParamsNonCopyable strParams;
        strParams.setup(configMap);

        const int device = pick_device(strParams.gpu);
        strParams.gpu = true;
        Queue Q(device);
        std::cout << "Using device: " << Q.device_name() << "\n";

        strParams.print();
        const Conv1dParams params = strParams;
        const size_t c_in = params.channel_in, c_out = params.channel_out, length = params.length;
        const size_t n0 = params.n0, n1 = params.n1, n2 = params.n2, k = params.k;

        span3d_t data(sycl_alloc(n0 * n1 * n2, Q), n0, n1, n2);
        span3d_t warmup_data(sycl_alloc(n0 * n1 * n2, Q), n0, n1, n2);
        span3d_t weight(sycl_alloc(k * c_out * c_in, Q), k, c_in, c_out);
        span1d_t bias{sycl_alloc(c_out, Q), c_out};
        fill_buffer_conv1d(Q, data, warmup_data, weight, bias);

        ConvSolver solver{weight, bias, k, c_in, length};

        BkmaOptimParams optim_params;
        throw_on(bkma_create_optim_params(n0, n1, n2, params.pref_wg_size, params.seq_size0, params.seq_size2,
                                          sizeof(real_t), Q.local_mem_size(), &optim_params));

        const double before = sum_and_normalize_conv1d(Q, data, n1);
        std::cout << "\nNormalized Array before: " << before << std::endl;

        for (int i = 0; i < 3; ++i) bkma_run<ConvSolver, BkmaImpl::AdaptiveWg>(Q, warmup_data, solver, optim_params).wait();

        const auto start = std::chrono::high_resolution_clock::now();
        bkma_run<ConvSolver, BkmaImpl::AdaptiveWg>(Q, data, solver, optim_params).wait();
        const auto end = std::chrono::high_resolution_clock::now();
        const std::chrono::duration<double> elapsed_seconds = end - start;

        const double after = sum_and_normalize_conv1d(Q, data, params.n_write);
        std::cout << "Normalized Array after: " << after << "\n" << std::endl;
        validate_conv1d(Q, data, params.n_write);

        std::cout << "PERF_DIAGS:" << std::endl;
        std::cout << "elapsed_time: " << elapsed_seconds.count() << " s\n";
        const double gcells = ((double)(n0 * n1 * n2) / elapsed_seconds.count()) / 1e9;
        std::cout << "upd_cells_per_sec: " << gcells << " Gcell/sec\n";
        std::cout << "estimated_throughput: " << gcells * sizeof(real_t) * 2 << " GB/s" << std::endl;

        device_free(weight.data_handle());
        device_free(bias.data_handle());
        device_free(data.data_handle());
        device_free(warmup_data.data_handle());
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
