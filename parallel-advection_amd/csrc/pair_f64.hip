// Compiled column-pair variants for FP64 storage (the tuner's search space).
// Each entry instantiates steps-per-pass 1 .. NSMAX of adv_column_pair_kernel
// (adv_pair.hpp).  The one-column-per-lane variant is compiled in
// pair_f64_v1.hip, so the two build in parallel.
#include "adv_pair.hpp"

namespace bkma {

const StreamVariant &pair_f64_vec1_variant();

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, DPP broadcast, NSMAX)
const StreamVariant *pair_variants_f64(int *count) {
    static const StreamVariant table[] = {
#ifdef BKMA_TUNED_FEW
        BKMA_PAIR_VARIANT(double, double, BKMA_F64, 2, 3, 14, 1, true, false, 4),
#else
        BKMA_PAIR_VARIANT(double, double, BKMA_F64, 2, 3, 14, 1, true, false, 5),
        pair_f64_vec1_variant(),
#endif
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace bkma
