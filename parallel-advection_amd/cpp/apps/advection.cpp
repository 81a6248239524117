// advection [file.ini]: the reference's advection driver (src/advection.cpp:14-71)
// on the gfx950 BKMA library.  kernelImpl = NDRange | AdaptiveWg selects the
// executor through impl_selector, which rejects every other name with the
// reference's message; kernelImpl = Tuned runs the tuned plan instead (FP64
// storage, bit-identical results).
#include <chrono>
#include <iostream>
#include <memory>

#include "bkma/bkma.hpp"
#include "bkma/tools.hpp"

using namespace bkma;

int main(int argc, char **argv) {
    try {
        const std::string input_file = argc > 1 ? std::string(argv[1]) : "advection.ini";
        ConfigMap configMap(input_file);

        ADVParamsNonCopyable strParams;
        strParams.setup(configMap);

        const int device = pick_device(strParams.gpu);
        strParams.gpu = 1;
        Queue Q(device);
        std::cout << "Using device: " << Q.device_name() << "\n";

        strParams.print();
        ADVParams params(strParams);
        const size_t n0 = params.n0, n1 = params.n1, n2 = params.n2, maxIter = params.maxIter;

        span3d_t data(sycl_alloc(n0 * n1 * n2, Q), n0, n1, n2);
        fill_buffer_adv(Q, data, params);

        const bool tuned = to_lowercase(strParams.kernelImpl) == "tuned";
        AdvectionSolver solver(params);
        span3d_t scratch;
        auto optim_params = create_optim_params<ADVParams>(Q, params);
        std::unique_ptr<TunedAdvection> plan;
        BkmaRunFn<AdvectionSolver> bkma_run_function;
        if (tuned) {
            plan = std::make_unique<TunedAdvection>(Q, params, BKMA_F64, true);
            std::cout << "Tuned plan: " << plan->name() << "\n";
        } else {
            bkma_run_function = impl_selector<AdvectionSolver>(strParams.kernelImpl);
            // A line longer than the LDS gets the global-scratch dispatch
            // (bkma_create_optim_params); AdaptiveWg then needs a caller span.
            if (to_lowercase(strParams.kernelImpl) == "adaptivewg" && optim_params.mem_space == BKMA_MEM_GLOBAL) {
                uint64_t e0, e1, e2;
                throw_on(bkma_scratch_extents(BKMA_IMPL_ADAPTIVE_WG, &optim_params, n0, n1, n2, 1, &e0, &e1, &e2));
                scratch = span3d_t(device_alloc<real_t>(e0 * e1 * e2), e0, e1, e2);
            }
        }

        const auto start = std::chrono::high_resolution_clock::now();
        if (tuned) {
            plan->run(Q, data, maxIter);  // the plan may fuse several steps per pass over HBM
            Q.wait();
        } else {
            for (size_t t = 0; t < maxIter; ++t) {
                bkma_run_function(Q, data, solver, optim_params, scratch);
                Q.wait();
            }
        }
        const auto end = std::chrono::high_resolution_clock::now();
        const std::chrono::duration<double> elapsed_seconds = end - start;

        validate_result_adv(Q, data, params);
        print_perf(elapsed_seconds.count(), n0 * n1 * n2 * maxIter);

        device_free(scratch.data_handle());
        device_free(data.data_handle());
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
