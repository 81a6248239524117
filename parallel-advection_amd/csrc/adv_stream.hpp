// Column-stream advection kernel: device code, launchers and the variant
// table type (the tables are instantiated in stream_f32.hip / stream_f64.hip,
// the plans in bkma_tuned.hip).
//
// One wavefront owns 64*VEC consecutive i2 columns of one i0 slab and walks
// the line in increasing i1 ("positions").  Each lane keeps, in registers, a
// window of WS = 2*SMAX+6 consecutive rows per stage: stage 0 holds rows
// read from memory, stage t (1 <= t < NS) holds the rows after t time steps
// (FP64), and stage NS is written back.  NS time steps are thus fused into
// one pass over HBM (one read and one write per cell for NS steps).
//
// Periodic wrap and in-place safety.  Positions run 0 .. R+n1+H-1 with
// R = NS*BACK and H = NS*AHEAD: positions [0, R) are rows n1-R .. n1-1,
// positions [R, R+n1) are rows 0 .. n1-1 and positions [R+n1, R+n1+H) are
// rows 0 .. H-1 again.  Stage t at position p produces the value of the
// row at p - t*AHEAD after t steps; only positions [R, R+n1) of the last
// stage are stored, each once, and every row is read from HBM before it is
// written.  Rows 0 .. H-1 are written before their second use, so their
// original values are kept in LDS.
//
// Foot location (displ, floor, d_prev1) and lag_basis are FP64 and in the
// reference's operation order.  They depend only on (i0, i1); each trip of
// U positions computes them once per position, one position per lane, and
// broadcasts them with readlane (identical bits to a per-cell evaluation).
// FP64 storage accumulates with separately rounded multiply and add as the
// reference does, and every intermediate stage holds exactly the FP64 value
// the reference stores after that step, so NS-step passes are bit-identical
// to NS single steps.  FP32 storage accumulates in FP64 with fused
// multiply-adds, keeps intermediate steps in FP64 and rounds to FP32 once
// per pass.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace bkma {

// Native vector types for 4/8/16-byte global accesses.
template <class T, int VEC> struct NativeVec { using type = T __attribute__((ext_vector_type(VEC))); };
template <class T> struct NativeVec<T, 1> { using type = T; };

template <class T, int VEC, bool NT> struct VecMem {
    using V = typename NativeVec<T, VEC>::type;
    __device__ __forceinline__ static void load(const T *p, T (&out)[VEC]) {
        V v;
        if constexpr (NT) v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
        else v = *reinterpret_cast<const V *>(p);
        if constexpr (VEC == 1) {
            out[0] = v;
        } else {
#pragma unroll
            for (int i = 0; i < VEC; ++i) out[i] = v[i];
        }
    }
    __device__ __forceinline__ static void store(T *p, const T (&in)[VEC]) {
        V v;
        if constexpr (VEC == 1) {
            v = in[0];
        } else {
#pragma unroll
            for (int i = 0; i < VEC; ++i) v[i] = in[i];
        }
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
        else *reinterpret_cast<V *>(p) = v;
    }
};

// Broadcast of a value every lane already holds identically, so that the
// compiler treats it (and branches on it) as wave-uniform.
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double wave_uniform(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Lane L's value (L a compile-time lane index), as a wave-uniform scalar.
template <int L> __device__ __forceinline__ int lane_value(int x) { return __builtin_amdgcn_readlane(x, L); }
template <int L> __device__ __forceinline__ double lane_value(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, L);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), L);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Shape of the per-stage window for shifts d = i1 - leftNode in [-SMAX, SMAX]:
// the stencil of the row at position q spans positions q-d-2 .. q-d+3.
template <int SMAX, int P, int NS> struct StreamShape {
    static constexpr int DLO = -SMAX, DHI = SMAX;
    static constexpr int BACK = SMAX + 2;   // positions behind the output position
    static constexpr int AHEAD = SMAX + 3;  // positions in front of it (stage lag)
    static constexpr int WS = BACK + AHEAD + 1;
    static constexpr int U = clcm(WS, P);  // unrolled positions per loop trip
    static constexpr int R = NS * BACK;    // positions before row 0
    static constexpr int H = NS * AHEAD;   // rows re-read from LDS at the end
    static constexpr int TB = NS * AHEAD;  // foot-table lane of offset 0
    static_assert(U + TB <= 64, "foot table must fit one wavefront");
};

// 6-tap interpolation of VEC cells from window slots S0..S0+5 (mod WS), in
// FP64.  FP64 storage (sizeof(T) == 8): separately rounded multiply and add;
// FP32 storage: fused multiply-add.
template <class T, class WT, class OT, int VEC, int WS, int S0>
__device__ __forceinline__ void interp_window(const WT (&win)[WS][VEC], const double (&c)[6], OT (&out)[VEC]) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        double acc = 0.;
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc = acc + c[k] * (double)win[(S0 + k) % WS][v];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc = fma(c[k], (double)win[(S0 + k) % WS][v], acc);
        }
        out[v] = (OT)acc;
    }
}

// Interpolation of the value at position offset OFF (mod WS) for the
// wave-uniform shift d: input slots OFF-d-2 .. OFF-d+3.
template <class T, class WT, class OT, int VEC, int WS, int OFF, int D, int DHI>
__device__ __forceinline__ void interp_shift(int d, const WT (&win)[WS][VEC], const double (&c)[6],
                                             OT (&out)[VEC]) {
    if constexpr (D > DHI) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) out[v] = (OT)NAN;  // outside the planned window: poison, never read memory
    } else {
        if (d == D) {
            interp_window<T, WT, OT, VEC, WS, cmod(OFF - D - 2, WS)>(win, c, out);
        } else {
            interp_shift<T, WT, OT, VEC, WS, OFF, D + 1, DHI>(d, win, c, out);
        }
    }
}

struct StreamArgs {
    AdvDev p;
    i64 tiles_per_slab;   // column tiles of 64*VEC per i0 slab
    i64 total_units;      // n0 * tiles_per_slab
    i64 wave_stride;      // waves in the grid (persistent loop stride)
    const int *slab_shift;  // column-pair kernels: smallest and largest shift of each i0 slab, [n0][2]
    int pair_ahead;         // column-pair kernels: head rows per fused step (3 - smallest shift)
};

// One wavefront advances unit (i0, column tile) by NS steps.
// WT: stage-0 window element type -- the storage type (widened at each use)
// or double (widened once, on entry).
template <class T, class WT, int VEC, int SMAX, int P, int NS, bool NT>
__device__ __forceinline__ void stream_unit(T *__restrict__ data, const StreamArgs &a, i64 unit, int lane,
                                            T (*head)[64 * VEC]) {
    using S = StreamShape<SMAX, P, NS>;
    constexpr int AHEAD = S::AHEAD, WS = S::WS, U = S::U, R = S::R, H = S::H, TB = S::TB;
    constexpr int NM = NS > 1 ? NS - 1 : 1;
    using M = VecMem<T, VEC, NT>;
    const AdvDev &p = a.p;
    const i64 n1 = p.n1, n2 = p.n2;
    const i64 i0 = unit / a.tiles_per_slab;
    const i64 tile = unit - i0 * a.tiles_per_slab;
    const i64 col = (tile * 64 + lane) * VEC;
    const bool lane_ok = col < n2;  // VEC divides n2: the whole vector is in range
    T *slab = data + (i0 * n1) * n2 + (lane_ok ? col : 0);
    const i64 first_end = R + n1;  // end of the positions read from HBM
    const i64 total = first_end + H;

    WT w0[WS][VEC];
    double wm[NM][WS][VEC];
    T stg[P][VEC];
#pragma unroll
    for (int s = 0; s < WS; ++s)
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            w0[s][v] = (WT)0;
#pragma unroll
            for (int t = 0; t < NM; ++t) wm[t][s][v] = 0.;
        }

    auto load_pos = [&](i64 pos, T (&dst)[VEC]) __attribute__((always_inline)) {
        const i64 row = pos < R ? pos - R + n1 : pos - R;  // pos < first_end
        if (lane_ok) {
            M::load(slab + row * n2, dst);
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) dst[v] = (T)0;
        }
    };
    static_for<0, P>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        load_pos(s, stg[s]);
    });

    const int half = (int)(n1 / 2);
    int tab_d = 0;
    double tab_c[6];

    for (i64 base = 0; base < total; base += U) {
        // Foot table: lane l describes the row at position base - TB + l.
        {
            i64 row = base - TB + lane - R;
            if (row < 0) row += n1;
            if (row >= n1) row -= n1;
            if (row < 0 || row >= n1) row = 0;  // positions no stored value depends on
            double d_prev1;
            const int leftNode = adv_foot(p, (int)i0, (int)row, d_prev1);
            adv_lag_basis(d_prev1, tab_c);
            int d = (int)row - leftNode;
            if (d > half) d -= (int)n1;
            if (d < -half) d += (int)n1;
            tab_d = d;
        }

        auto step = [&](auto jc) __attribute__((always_inline)) {
            constexpr int J = decltype(jc)::value;
            const i64 pin = base + J;
            if (pin >= total) return;
            // 1. position pin enters the stage-0 window
            {
                constexpr int sw = J % WS, ss = J % P;
                if (pin < first_end) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v) w0[sw][v] = (WT)stg[ss][v];
                    if (pin >= R && pin < R + H) {
#pragma unroll
                        for (int v = 0; v < VEC; ++v) head[pin - R][lane * VEC + v] = stg[ss][v];
                    }
                    // 2. refill the staging slot with position pin + P
                    if (pin + P < first_end) load_pos(pin + P, stg[ss]);
                } else {
                    const int h = (int)(pin - first_end);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) w0[sw][v] = (WT)head[h][lane * VEC + v];
                }
            }
            // 3. stage t computes position pin - t*AHEAD from stage t-1
            static_for<1, NS + 1>([&](auto tc) __attribute__((always_inline)) {
                constexpr int t = decltype(tc)::value;
                constexpr int OFF = J - t * AHEAD;
                constexpr int L = OFF + TB;
                const int d = lane_value<L>(tab_d);
                double c[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) c[k] = lane_value<L>(tab_c[k]);
                if constexpr (t < NS) {
                    double o[VEC];
                    if constexpr (t == 1)
                        interp_shift<T, WT, double, VEC, WS, OFF, S::DLO, S::DHI>(d, w0, c, o);
                    else
                        interp_shift<T, double, double, VEC, WS, OFF, S::DLO, S::DHI>(d, wm[t - 2], c, o);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) wm[t - 1][cmod(OFF, WS)][v] = o[v];
                } else {
                    T o[VEC];
                    if constexpr (NS == 1)
                        interp_shift<T, WT, T, VEC, WS, OFF, S::DLO, S::DHI>(d, w0, c, o);
                    else
                        interp_shift<T, double, T, VEC, WS, OFF, S::DLO, S::DHI>(d, wm[NS - 2], c, o);
                    const i64 pt = pin - NS * AHEAD;
                    if (pt >= R && pt < first_end && lane_ok) M::store(slab + (pt - R) * n2, o);
                }
            });
        };
        static_for<0, U>(step);
    }
}

template <class T, class WT, int VEC, int SMAX, int P, int NS, int WPB, bool NT>
__global__ __launch_bounds__(64 * WPB) void adv_column_stream_kernel(T *__restrict__ data, StreamArgs a) {
    using S = StreamShape<SMAX, P, NS>;
    __shared__ T head_lds[WPB][S::H][64 * VEC];
    const int lane = threadIdx.x & 63;
    const int wv = wave_uniform((int)(threadIdx.x >> 6));  // unit indices and addresses stay scalar
    // Persistent waves: unit = wave, wave + stride, ...  (all lanes of a wave
    // take the same trip count; there is no work-group barrier).
    for (i64 unit = (i64)blockIdx.x * WPB + wv; unit < a.total_units; unit += a.wave_stride)
        stream_unit<T, WT, VEC, SMAX, P, NS, NT>(data, a, unit, lane, head_lds[wv]);
}

// ---------------------------------------------------------------------------
// Host side: launchers for steps-per-pass 1 .. NSMAX of one variant
// ---------------------------------------------------------------------------
using StreamLaunchFn = void (*)(void *data, const StreamArgs &a, unsigned grid, hipStream_t s, int ns);
using StreamOccFn = int (*)(int ns, int pair_ahead);

template <class T, class WT, int VEC, int SMAX, int P, int WPB, bool NT, int NS>
void launch_stream(void *data, const StreamArgs &a, unsigned grid, hipStream_t s, int ns) {
    if constexpr (NS >= 1) {
        if (ns == NS) {
            hipLaunchKernelGGL((adv_column_stream_kernel<T, WT, VEC, SMAX, P, NS, WPB, NT>), dim3(grid),
                               dim3(64 * WPB), 0, s, reinterpret_cast<T *>(data), a);
            return;
        }
        launch_stream<T, WT, VEC, SMAX, P, WPB, NT, NS - 1>(data, a, grid, s, ns);
    }
}

// Resident work-groups per CU (0 if the query fails or ns is not compiled).
template <class T, class WT, int VEC, int SMAX, int P, int WPB, bool NT, int NS>
int occupancy_stream(int ns, int pair_ahead = 0) {
    if constexpr (NS >= 1) {
        if (ns == NS) {
            int blocks = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &blocks, adv_column_stream_kernel<T, WT, VEC, SMAX, P, NS, WPB, NT>, 64 * WPB, 0) != hipSuccess)
                return 0;
            return blocks;
        }
        return occupancy_stream<T, WT, VEC, SMAX, P, WPB, NT, NS - 1>(ns, pair_ahead);
    }
    return 0;
}

struct StreamVariant {
    bkma_dtype dtype;
    int pair;                                // 0: shift ring (this file), 1: per-slab shift pair (adv_pair.hpp)
    int vec, smax, prefetch, wpb, nt, wide;  // wide: FP64 stage-0 window for FP32 storage
    int nsmax;                               // steps per pass compiled: 1 .. nsmax
    StreamLaunchFn fn;
    StreamOccFn occ;
    int dpp;  // column pair: coefficients broadcast by DPP (1) or v_readlane (0)
};

#define BKMA_STREAM_VARIANT(T, WT, DT, V, S, P, W, NT, NSMAX)                                                  \
    StreamVariant {                                                                                           \
        DT, 0, V, S, P, W, NT, sizeof(WT) > sizeof(T), NSMAX, &launch_stream<T, WT, V, S, P, W, NT, NSMAX>,   \
            &occupancy_stream<T, WT, V, S, P, W, NT, NSMAX>, 0                                                \
    }

// The compiled variants of each storage type (host functions): shift-ring
// kernels (stream_f*.hip) and column-pair kernels (pair_f*.hip).
const StreamVariant *stream_variants_f32(int *count);
const StreamVariant *stream_variants_f64(int *count);
const StreamVariant *pair_variants_f32(int *count);
const StreamVariant *pair_variants_f64(int *count);

}  // namespace bkma
