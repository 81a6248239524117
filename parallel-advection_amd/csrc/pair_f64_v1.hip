// Column-pair variant for FP64 storage with one column per lane: the window
// of more fused steps fits (195 VGPRs at 8).  Listed in pair_f64.hip.
#include "adv_pair.hpp"

namespace bkma {

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, DPP broadcast, NSMAX)
const StreamVariant &pair_f64_vec1_variant() {
    static const StreamVariant v = BKMA_PAIR_VARIANT(double, double, BKMA_F64, 1, 3, 14, 1, true, false, 8);
    return v;
}

}  // namespace bkma
