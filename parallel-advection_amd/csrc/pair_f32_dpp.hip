// Column-pair variant for FP32 storage with the DPP coefficient broadcast
// (the headline kernel, up to 5 fused steps).  Listed in pair_f32.hip.
#include "adv_pair.hpp"

namespace bkma {

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, DPP broadcast, NSMAX)
const StreamVariant &pair_f32_dpp_variant() {
    static const StreamVariant v = BKMA_PAIR_VARIANT(float, double, BKMA_F32, 2, 3, 14, 1, true, true, 5);
    return v;
}

}  // namespace bkma
