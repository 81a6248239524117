// Compiled column-stream variants for FP32 storage (the tuner's search space).
// Each entry instantiates steps-per-pass 1 .. NSMAX of the shift-ring kernel
// (adv_stream.hpp); the column-pair variants are in pair_f32.hip.
#include "adv_stream.hpp"

namespace bkma {

// (storage, stage-0 window, dtype, VEC, SMAX, P, waves/WG, non-temporal, NSMAX)
// P is WS/2 or WS (WS = 2*SMAX + 6) so that one loop trip is one window long.
const StreamVariant *stream_variants_f32(int *count) {
    static const StreamVariant table[] = {
#ifdef BKMA_TUNED_FEW
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 4, 2, 10, 1, true, 2),
#else
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 4, 2, 10, 1, true, 3),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 4, 2, 5, 1, true, 3),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 2, 2, 10, 1, true, 4),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 1, 2, 10, 1, false, 4),
        BKMA_STREAM_VARIANT(float, double, BKMA_F32, 4, 2, 10, 1, true, 1),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 4, 1, 8, 1, true, 3),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 2, 1, 8, 1, true, 4),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 4, 4, 7, 1, true, 2),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 2, 4, 14, 1, true, 3),
        BKMA_STREAM_VARIANT(float, float, BKMA_F32, 1, 4, 14, 1, false, 3),
#endif
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace bkma

