// Contiguous lines (n2 = 1) advanced in registers, stencil neighbours
// through wavefront shuffles (strategy LINE_SHFL, see bkma_tuned.hip).
//
// One work-group owns one line of n1 = 64 * E * NW cells (E = 4 or 8 cells
// per lane): lane g of the work-group holds cells [g*E, g*E + E) in FP64
// registers for a whole pass of ns steps.  The stencil of cell i reads cells
// i+rel .. i+rel+5 with rel = leftNode - 2 - i.  On a line whose rows all
// take one shift, rel is one value DMIN and every cell sums a six-cell window
// from offset DMIN; on a line whose rows take two adjacent shifts (at most
// two, checked at plan creation) rel is DMIN or DMIN+1 and every cell sums a
// seven-cell window with seven cached coefficients, one of them zero.  With
// DMIN in [-E, E-TAPS+1] the window lies in the lane's own cells and those of
// lanes g-1 and g+1, which arrive through DPP:
//  - wave_shr:1 / wave_shl:1 move a register one lane up / down the wave;
//    lanes 0 and 63 keep the DPP `old` operand, which holds the neighbouring
//    wave's edge cells, exchanged through a small LDS record each step;
//  - a line held by one wave wraps with wave_ror:1 / wave_rol:1 instead.
// The coefficients depend only on (i0, i1), so they are computed once per
// pass (the reference's foot location and lag_basis, in FP64) and reused by
// every step of it.
//
// FP64 storage: the reference's separately rounded products and sums in its
// order (the zero term adds an exact zero), bit-identical for any steps per
// pass.  FP32 storage: fused FP64 accumulation, FP64 values between the steps
// of a pass, one rounding to FP32 per pass.
#pragma once

#include "common.hpp"

namespace bkma {

constexpr int kShflMaxSteps = 256;  // steps per pass

// Window start range of a lane holding E cells with TAPS-cell windows: the
// window of its last cell must end in lane g+1, that of its first start in
// lane g-1.
template <int E, int TAPS> constexpr int shfl_dmin_lo() { return -E; }
template <int E, int TAPS> constexpr int shfl_dmin_hi() { return E - TAPS + 1; }

struct ShflArgs {
    AdvDev p;
    const int *slab_shift;  // [n0][2]: smallest and largest d = i1 - leftNode of each line
    int ns;                 // steps in this pass
};

// DPP on the two halves of an FP64 register.  Lanes without a source lane
// keep `old` (bound_ctrl off).
template <int CTRL> __device__ __forceinline__ double dpp_f64(double old, double src) {
    const long long o = __double_as_longlong(old), s = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// DPP move of an FP64 register where every lane has a source lane
// (rotations): no `old` operand, so no copy of it into the destination.
template <int CTRL> __device__ __forceinline__ double dpp_rot_f64(double src) {
    const long long v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
constexpr int kDppWaveShl1 = 0x130;  // lane l reads lane l+1
constexpr int kDppWaveRol1 = 0x134;  // lane l reads lane (l+1) mod 64
constexpr int kDppWaveShr1 = 0x138;  // lane l reads lane l-1
constexpr int kDppWaveRor1 = 0x13C;  // lane l reads lane (l-1) mod 64

// One step: g from f (this lane's E cells), windows of TAPS cells starting
// at offset DMIN; buf is this step's half of the LDS exchange record (NW
// waves x {lane 0's cells, lane 63's cells}).
template <class T, int E, int NW, int TAPS, int DMIN>
__device__ __forceinline__ void shfl_step(const double (&f)[E], double (&g)[E], const double (&c)[E][TAPS],
                                          double *buf, int w, int wl, int wr, int lane) {
    constexpr int NL = DMIN < 0 ? -DMIN : 0;                     // cells from lane g-1: its slots E-NL .. E-1
    constexpr int NR = DMIN + TAPS - 1 > 0 ? DMIN + TAPS - 1 : 0;  // cells from lane g+1: its slots 0 .. NR-1
    static_assert(DMIN >= shfl_dmin_lo<E, TAPS>() && DMIN <= shfl_dmin_hi<E, TAPS>(),
                  "window outside lanes g-1 .. g+1");
    double hl[NL > 0 ? NL : 1], hr[NR > 0 ? NR : 1];
    if constexpr (NW > 1) {
        // edge cells of the neighbouring waves (periodic over the line)
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < NR; ++j) buf[(2 * w) * E + j] = f[j];
        }
        if (lane == 63) {
#pragma unroll
            for (int j = 0; j < NL; ++j) buf[(2 * w + 1) * E + E - NL + j] = f[E - NL + j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NL; ++j) hl[j] = buf[(2 * wl + 1) * E + E - NL + j];
#pragma unroll
        for (int j = 0; j < NR; ++j) hr[j] = buf[(2 * wr) * E + j];
    }
    // W[t] = the cell at offset DMIN + t from this lane's first cell
    double W[E + TAPS - 1];
    static_for<0, E + TAPS - 1>([&](auto tt) {
        constexpr int t = decltype(tt)::value;
        constexpr int o = DMIN + t;
        if constexpr (o < 0) {
            if constexpr (NW > 1) W[t] = dpp_f64<kDppWaveShr1>(hl[o + NL], f[E + o]);
            else W[t] = dpp_rot_f64<kDppWaveRor1>(f[E + o]);
        } else if constexpr (o < E) {
            W[t] = f[o];
        } else {
            if constexpr (NW > 1) W[t] = dpp_f64<kDppWaveShl1>(hr[o - E], f[o - E]);
            else W[t] = dpp_rot_f64<kDppWaveRol1>(f[o - E]);
        }
    });
#pragma unroll
    for (int e = 0; e < E; ++e) {
        double acc;
        if constexpr (sizeof(T) == 8) {
            acc = c[e][0] * W[e];
#pragma unroll
            for (int t = 1; t < TAPS; ++t) acc = acc + c[e][t] * W[e + t];
        } else {
            acc = 0.;
#pragma unroll
            for (int t = 0; t < TAPS; ++t) acc = fma(c[e][t], W[e + t], acc);
        }
        g[e] = acc;
    }
}

// ns steps of one line held in f.  Steps alternate between two register
// sets (f -> g -> f), so no step copies its results back; the LDS exchange
// record alternates halves with the step parity (one barrier per step).
template <class T, int E, int NW, int TAPS, int DMIN>
__device__ __forceinline__ void shfl_steps(double (&f)[E], const double (&c)[E][TAPS], int ns, double *xch, int w,
                                           int lane) {
    const int wl = w == 0 ? NW - 1 : w - 1, wr = w == NW - 1 ? 0 : w + 1;
    double *buf1 = xch + NW * 2 * E;
    double g[E];
    int s = 0;
    for (; s + 2 <= ns; s += 2) {
        shfl_step<T, E, NW, TAPS, DMIN>(f, g, c, xch, w, wl, wr, lane);
        shfl_step<T, E, NW, TAPS, DMIN>(g, f, c, buf1, w, wl, wr, lane);
    }
    if (s < ns) {
        shfl_step<T, E, NW, TAPS, DMIN>(f, g, c, xch, w, wl, wr, lane);
#pragma unroll
        for (int e = 0; e < E; ++e) f[e] = g[e];
    }
}

template <class T, int E, int NW, int TAPS, int D>
__device__ __forceinline__ void shfl_dispatch(int dmin, double (&f)[E], const double (&c)[E][TAPS], int ns,
                                              double *xch, int w, int lane) {
    if constexpr (D <= shfl_dmin_hi<E, TAPS>()) {
        if (dmin == D) shfl_steps<T, E, NW, TAPS, D>(f, c, ns, xch, w, lane);
        else shfl_dispatch<T, E, NW, TAPS, D + 1>(dmin, f, c, ns, xch, w, lane);
    }
}

// Coefficients of this lane's E cells over windows of TAPS cells from the
// line's window start (TAPS = 6: every row of the line has the shift dmax;
// TAPS = 7: rows with the shift dmax - 1 start one cell later, a zero first),
// then the steps.
template <class T, int E, int NW, int TAPS>
__device__ __forceinline__ void shfl_line(double (&f)[E], const AdvDev &p, i64 i0, int g, int dmax, int ns,
                                          double *xch, int w, int lane) {
    const i64 n1 = p.n1;
    const int half = (int)(n1 / 2);
    double c[E][TAPS];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i1 = g * E + e;
        double d_prev1;
        const int leftNode = adv_foot(p, (int)i0, i1, d_prev1);
        double cf[6];
        adv_lag_basis(d_prev1, cf);
        if constexpr (TAPS == 6) {
#pragma unroll
            for (int t = 0; t < 6; ++t) c[e][t] = cf[t];
        } else {
            int d = i1 - leftNode;
            if (d > half) d -= (int)n1;
            if (d < -half) d += (int)n1;
            const bool up = d < dmax;  // rel = dmin + 1: the taps start one cell later
#pragma unroll
            for (int t = 0; t < 7; ++t) {
                const double lo = t < 6 ? cf[t] : 0.0, hi = t > 0 ? cf[t - 1] : 0.0;
                c[e][t] = up ? hi : lo;
            }
        }
    }
    shfl_dispatch<T, E, NW, TAPS, shfl_dmin_lo<E, TAPS>()>(-dmax - 2, f, c, ns, xch, w, lane);
}

template <class T, int E, int NW>
__global__ __launch_bounds__(64 * NW) void adv_line_shfl_kernel(T *__restrict__ data, ShflArgs a) {
    static_assert(E == 4 || E == 8, "4 or 8 cells per lane");
    __shared__ double xch[NW > 1 ? 2 * NW * 2 * E : 1];
    const AdvDev &p = a.p;
    const i64 i0 = blockIdx.x;
    const int g = threadIdx.x, lane = g & 63, w = g >> 6;
    T *line = data + i0 * p.n1 + (i64)g * E;  // n1 = 64 * E * NW

    double f[E];
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < E / 4; ++q) {
            const float4 v = reinterpret_cast<const float4 *>(line)[q];
            f[4 * q] = v.x, f[4 * q + 1] = v.y, f[4 * q + 2] = v.z, f[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < E / 2; ++q) {
            const double2 v = reinterpret_cast<const double2 *>(line)[q];
            f[2 * q] = v.x, f[2 * q + 1] = v.y;
        }
    }
    // this line's shifts d = i1 - leftNode; the window starts at -dmax - 2
    const int dmin_row = a.slab_shift[2 * i0], dmax = a.slab_shift[2 * i0 + 1];
    if (dmin_row == dmax) shfl_line<T, E, NW, 6>(f, p, i0, g, dmax, a.ns, xch, w, lane);
    else shfl_line<T, E, NW, 7>(f, p, i0, g, dmax, a.ns, xch, w, lane);
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < E / 4; ++q)
            reinterpret_cast<float4 *>(line)[q] =
                make_float4((float)f[4 * q], (float)f[4 * q + 1], (float)f[4 * q + 2], (float)f[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < E / 2; ++q) reinterpret_cast<double2 *>(line)[q] = make_double2(f[2 * q], f[2 * q + 1]);
    }
}

// Lines of 16 cells: one line per 16-lane DPP row, one cell per lane, four
// lines per wavefront.  Every cell's window covers the same offsets
// WMIN .. WMIN+TW-1 (the union of the windows of all lines, TW <= 8); a
// cell's six coefficients sit at its own offset in it and the other taps are
// zero.  Neighbour o of lane r is lane (r + o) mod 16 of the row: one DPP
// row_ror:((16 - o) mod 16) per tap and FP64 half.
template <int O> constexpr int row_ror_ctrl() { return 0x120 + ((16 - (O % 16)) % 16); }

template <class T, int WMIN, int TW>
__global__ __launch_bounds__(256) void adv_line_row16_kernel(T *__restrict__ data, ShflArgs a) {
    static_assert(TW >= 6 && TW <= 8, "window of 6 to 8 cells");
    const AdvDev &p = a.p;
    const int r = threadIdx.x & 15;
    const i64 i0 = (i64)blockIdx.x * 16 + (threadIdx.x >> 4);
    const bool valid = i0 < p.n0;  // whole rows are valid or not
    T *cell = data + (valid ? i0 : 0) * 16 + r;
    double f = valid ? (double)*cell : 0.0;

    double c[TW];
    {
        double d_prev1;
        const int leftNode = valid ? adv_foot(p, (int)i0, r, d_prev1) : r;
        if (!valid) d_prev1 = 2.0;
        double cf[6];
        adv_lag_basis(d_prev1, cf);
        int rel = leftNode - 2 - r;  // first tap's offset, normalised into (-8, 8]
        if (rel > 8) rel -= 16;
        if (rel <= -8) rel += 16;
        const int k = valid ? rel - WMIN : 0;  // in [0, TW - 6] (checked at plan creation)
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            double v = 0.0;
#pragma unroll
            for (int j = 0; j <= TW - 6; ++j)
                if (t - j >= 0 && t - j < 6 && k == j) v = cf[t - j];
            c[t] = v;
        }
    }
    for (int s = 0; s < a.ns; ++s) {
        double W[TW];
        static_for<0, TW>([&](auto tt) {
            constexpr int t = decltype(tt)::value;
            constexpr int o = WMIN + t;
            if constexpr (o % 16 == 0) W[t] = f;
            else W[t] = dpp_rot_f64<row_ror_ctrl<o>()>(f);
        });
        double acc;
        if constexpr (sizeof(T) == 8) {
            acc = c[0] * W[0];
#pragma unroll
            for (int t = 1; t < TW; ++t) acc = acc + c[t] * W[t];
        } else {
            acc = 0.;
#pragma unroll
            for (int t = 0; t < TW; ++t) acc = fma(c[t], W[t], acc);
        }
        f = acc;
    }
    if (valid) *cell = (T)f;
}

// Host side (line_shfl.hip): one pass of a.ns steps over n0 lines of
// 64 * cells * nw cells; cells 4 with nw in 1 .. 16, cells 8 with nw in
// 1 .. 8 (from 9 wavefronts per work-group on, a wave has at most 128 VGPRs,
// fewer than 8 cells' cached coefficients need).
bkma_status launch_line_shfl(bkma_dtype dt, int cells, int nw, void *data, const ShflArgs &a, i64 n0,
                             hipStream_t s);
// Lines of 16 cells (adv_line_row16_kernel): window offsets wmin .. wmin+tw-1
// with wmin in [-6, 0] and tw in [6, 8].
bkma_status launch_line_row16(bkma_dtype dt, int wmin, int tw, void *data, const ShflArgs &a, i64 n0, hipStream_t s);

}  // namespace bkma
