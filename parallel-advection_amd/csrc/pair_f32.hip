// Compiled column-pair variants for FP32 storage (the tuner's search space).
// Each entry instantiates steps-per-pass 1 .. NSMAX of adv_column_pair_kernel
// (adv_pair.hpp).
#include "adv_pair.hpp"

namespace bkma {

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, DPP broadcast, NSMAX)
const StreamVariant *pair_variants_f32(int *count) {
    static const StreamVariant table[] = {
#ifdef BKMA_TUNED_FEW
        BKMA_PAIR_VARIANT(float, double, BKMA_F32, 2, 3, 14, 1, true, true, 5),
#else
        BKMA_PAIR_VARIANT(float, double, BKMA_F32, 4, 3, 7, 1, true, false, 3),
        BKMA_PAIR_VARIANT(float, double, BKMA_F32, 2, 3, 14, 1, true, true, 5),
#endif
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace bkma
