// Instances and launcher of the register/DPP kernel for contiguous lines
// (adv_shfl.hpp).
#include "adv_shfl.hpp"

namespace bkma {

namespace {

template <class T, int E, int NW>
bkma_status launch_nw(void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    hipLaunchKernelGGL((adv_line_shfl_kernel<T, E, NW>), dim3((unsigned)n0), dim3(64 * NW), 0, s, (T *)data, a);
    return check_launch("adv_line_shfl_kernel");
}

// Wavefronts per line: 1 .. 16 with 4 cells per lane, 1 .. 8 with 8.
template <class T, int E, int NW>
bkma_status launch_nw_from(int nw, void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    if constexpr (NW > (E == 4 ? 16 : 8)) {
        return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: wavefronts per line not compiled");
    } else {
        if (nw == NW) return launch_nw<T, E, NW>(data, a, n0, s);
        return launch_nw_from<T, E, NW + 1>(nw, data, a, n0, s);
    }
}

template <class T, int E>
bkma_status launch_e(int nw, void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    return launch_nw_from<T, E, 1>(nw, data, a, n0, s);
}

template <class T>
bkma_status launch_t(int cells, int nw, void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    if (cells == 4) return launch_e<T, 4>(nw, data, a, n0, s);
    if (cells == 8) return launch_e<T, 8>(nw, data, a, n0, s);
    return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: 4 or 8 cells per lane");
}

template <class T, int WMIN, int TW>
bkma_status launch_row(void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    const i64 groups = (n0 + 15) / 16;
    hipLaunchKernelGGL((adv_line_row16_kernel<T, WMIN, TW>), dim3((unsigned)groups), dim3(256), 0, s, (T *)data, a);
    return check_launch("adv_line_row16_kernel");
}

template <class T, int WMIN>
bkma_status launch_row_w(int tw, void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    switch (tw) {
        case 6: return launch_row<T, WMIN, 6>(data, a, n0, s);
        case 7: return launch_row<T, WMIN, 7>(data, a, n0, s);
        case 8: return launch_row<T, WMIN, 8>(data, a, n0, s);
        default: return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL rows: window of 6 to 8 cells");
    }
}

template <class T>
bkma_status launch_row_t(int wmin, int tw, void *data, const ShflArgs &a, i64 n0, hipStream_t s) {
    switch (wmin) {
        case -6: return launch_row_w<T, -6>(tw, data, a, n0, s);
        case -5: return launch_row_w<T, -5>(tw, data, a, n0, s);
        case -4: return launch_row_w<T, -4>(tw, data, a, n0, s);
        case -3: return launch_row_w<T, -3>(tw, data, a, n0, s);
        case -2: return launch_row_w<T, -2>(tw, data, a, n0, s);
        case -1: return launch_row_w<T, -1>(tw, data, a, n0, s);
        case 0: return launch_row_w<T, 0>(tw, data, a, n0, s);
        default: return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL rows: window start out of range");
    }
}

}  // namespace

bkma_status launch_line_row16(bkma_dtype dt, int wmin, int tw, void *data, const ShflArgs &a, i64 n0,
                              hipStream_t s) {
    if (n0 <= 0 || (n0 + 15) / 16 > 0x7fffffffLL) return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: line count out of range");
    if (a.ns < 1 || a.ns > kShflMaxSteps) return fail(BKMA_ERR_RUNTIME, "LINE_SHFL: steps per pass out of range");
    if (a.p.n2 != 1 || a.p.n1 != 16) return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL rows: lines of 16 cells");
    if (dt == BKMA_F64) return launch_row_t<double>(wmin, tw, data, a, n0, s);
    if (dt == BKMA_F32) return launch_row_t<float>(wmin, tw, data, a, n0, s);
    return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: F32 or F64 storage");
}

bkma_status launch_line_shfl(bkma_dtype dt, int cells, int nw, void *data, const ShflArgs &a, i64 n0,
                             hipStream_t s) {
    if (n0 <= 0 || n0 > 0x7fffffffLL) return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: line count out of range");
    if (a.ns < 1 || a.ns > kShflMaxSteps) return fail(BKMA_ERR_RUNTIME, "LINE_SHFL: steps per pass out of range");
    if (a.p.n2 != 1 || a.p.n1 != (i64)64 * cells * nw)
        return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: line shape does not match the work-group");
    if (dt == BKMA_F64) return launch_t<double>(cells, nw, data, a, n0, s);
    if (dt == BKMA_F32) return launch_t<float>(cells, nw, data, a, n0, s);
    return fail(BKMA_ERR_UNSUPPORTED, "LINE_SHFL: F32 or F64 storage");
}

}  // namespace bkma
