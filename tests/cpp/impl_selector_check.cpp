// impl_selector of the C++ facade against the reference's names and error
// text (src/tools/impl_selector.hpp:9-10, 31-49): prints one line per name,
// "ok" or the exception's type and message.  No GPU work is done: the
// selector only returns a function.  TEST INFRASTRUCTURE (tests/test_facade.py).
#include <iostream>
#include <stdexcept>

#include "bkma/bkma.hpp"

int main() {
    const char *names[] = {"AdaptiveWg", "adaptivewg", "ADAPTIVEWG", "NDRange", "ndrange",
                           "BasicRange", "basicrange", "Tuned", "", "AdaptiveWg "};
    for (const char *n : names) {
        try {
            auto f = bkma::impl_selector<bkma::AdvectionSolver>(n);
            auto g = bkma::impl_selector<bkma::ConvSolver>(n);
            std::cout << "[" << n << "] ok " << (f ? 1 : 0) << (g ? 1 : 0) << "\n";
        } catch (const std::runtime_error &e) {
            std::cout << "[" << n << "] runtime_error <" << e.what() << ">\n";
        }
    }
    return 0;
}
