// Compiled column-pair variants for FP64 storage (the tuner's search space).
// Each entry instantiates steps-per-pass 1 .. NSMAX of adv_column_pair_kernel
// (adv_pair.hpp).
#include "adv_pair.hpp"

namespace bkma {

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, DPP broadcast, NSMAX)
const StreamVariant *pair_variants_f64(int *count) {
    static const StreamVariant table[] = {
#ifdef BKMA_TUNED_FEW
        BKMA_PAIR_VARIANT(double, double, BKMA_F64, 2, 3, 14, 1, true, false, 4),
#else
        BKMA_PAIR_VARIANT(double, double, BKMA_F64, 2, 3, 14, 1, true, false, 5),
        // one column per lane: the window of more fused steps fits (195 VGPRs at 8)
        BKMA_PAIR_VARIANT(double, double, BKMA_F64, 1, 3, 14, 1, true, false, 8),
#endif
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace bkma
