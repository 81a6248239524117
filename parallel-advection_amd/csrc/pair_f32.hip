// Compiled column-pair variants for FP32 storage (the tuner's search space).
// Each entry instantiates steps-per-pass 1 .. NSMAX of adv_column_pair_kernel
// (adv_pair.hpp).  The DPP-broadcast variant is compiled in
// pair_f32_dpp.hip, so the two build in parallel.
#include "adv_pair.hpp"

namespace bkma {

const StreamVariant &pair_f32_dpp_variant();

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, DPP broadcast, NSMAX)
const StreamVariant *pair_variants_f32(int *count) {
    static const StreamVariant table[] = {
#ifdef BKMA_TUNED_FEW
        pair_f32_dpp_variant(),
#else
        BKMA_PAIR_VARIANT(float, double, BKMA_F32, 4, 3, 7, 1, true, false, 3),
        pair_f32_dpp_variant(),
#endif
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace bkma
