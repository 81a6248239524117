// Device side of the column-stream advection kernel (see bkma_tuned.hip).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace bkma {

template <int B, int E, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int clcm(int a, int b) { return a / cgcd(a, b) * b; }

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

// Shape of the stencil window for shifts d = i1 - leftNode in [-SMAX, SMAX].
template <int SMAX, int P> struct StreamShape {
    static constexpr int DLO = -SMAX, DHI = SMAX;
    static constexpr int BACK = SMAX + 2;   // rows behind the output row
    static constexpr int AHEAD = SMAX + 3;  // rows in front of it
    static constexpr int WS = BACK + AHEAD + 1;
    static constexpr int U = clcm(WS, P);   // unrolled rows per loop trip
};

// 6-tap interpolation of VEC cells from the window slots S0..S0+5 (mod WS),
// in FP64.  FP64 storage: the reference's separately rounded multiply and
// add; FP32 storage: fused FP64 multiply-add.
template <class T, class WT, int VEC, int WS, int S0>
__device__ __forceinline__ void interp_window(const WT (&win)[WS][VEC], const double (&c)[6], T (&out)[VEC]) {
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
        out[v] = (T)acc;
    }
}

// Select the window slots for the (wave-uniform) shift d at loop position J.
template <class T, class WT, int VEC, int WS, int J, int BACK, int D, int DHI>
__device__ __forceinline__ void interp_shift(int d, const WT (&win)[WS][VEC], const double (&c)[6],
                                             T (&out)[VEC]) {
    if constexpr (D > DHI) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) out[v] = (T)NAN;  // outside the planned window: poison, never read memory
    } else {
        if (d == D) {
            constexpr int S0 = (((J + BACK - D - 2) % WS) + WS) % WS;
            interp_window<T, WT, VEC, WS, S0>(win, c, out);
        } else {
            interp_shift<T, WT, VEC, WS, J, BACK, D + 1, DHI>(d, win, c, out);
        }
    }
}

struct StreamArgs {
    AdvDev p;
    i64 tiles_per_slab;  // column tiles of 64*VEC per i0 slab
    i64 total_units;     // n0 * tiles_per_slab
    i64 wave_stride;     // waves in the grid (persistent loop stride)
};

// One wavefront streams unit (i0, column tile) through the whole line.
// Row r of the line lives, in turn, in: the staging ring (loaded, P rows
// ahead), the register window (rows q-BACK .. q+AHEAD around output q).
// WT: window element type -- double (FP32 rows widened once, on entry) or
// the storage type (widened at each use; fewer registers).
template <class T, class WT, int VEC, int SMAX, int P, bool NT>
__device__ __forceinline__ void stream_unit(T *__restrict__ data, const StreamArgs &a, i64 unit, int lane,
                                            T (*head)[64 * VEC]) {
    using S = StreamShape<SMAX, P>;
    constexpr int BACK = S::BACK, AHEAD = S::AHEAD, WS = S::WS, U = S::U;
    using M = VecMem<T, VEC, NT>;
    const AdvDev &p = a.p;
    const i64 n1 = p.n1, n2 = p.n2;
    const i64 i0 = unit / a.tiles_per_slab;
    const i64 tile = unit - i0 * a.tiles_per_slab;
    const i64 col = (tile * 64 + lane) * VEC;
    const bool lane_ok = col < n2;  // VEC divides n2: the whole vector is in range
    T *slab = data + (i0 * n1) * n2 + (lane_ok ? col : 0);

    WT win[WS][VEC];
    T stg[P][VEC];
    auto load_row = [&](i64 row, T (&dst)[VEC]) __attribute__((always_inline)) {
        if (lane_ok) {
            M::load(slab + row * n2, dst);
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) dst[v] = (T)0;
        }
    };

    // Prologue: window rows [-BACK, AHEAD) (tail rows wrap to n1-BACK..n1-1),
    // rows [0, AHEAD) also kept for the wrap at the end of the line; staging
    // rows [AHEAD, AHEAD+P).
    static_for<0, BACK + AHEAD>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        constexpr int r = s - BACK;
        T tmp[VEC];
        load_row(r < 0 ? n1 + r : r, tmp);
#pragma unroll
        for (int v = 0; v < VEC; ++v) win[s][v] = (WT)tmp[v];
        if constexpr (r >= 0) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) head[r][lane * VEC + v] = tmp[v];
        }
    });
    static_for<0, P>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        load_row(AHEAD + s, stg[(AHEAD + s) % P]);
    });

    double coef[6];
    unsigned long long cached_bits = ~0ull;  // not the bits of any d_prev1 >= 0
    const int half = (int)(n1 / 2);
    i64 base = 0;

    auto row_step = [&](auto jc) __attribute__((always_inline)) {
        constexpr int J = decltype(jc)::value;
        const i64 q = base + J;
        if (q >= n1) return;
        // 1. row q+AHEAD enters the window (slot of row q-BACK-1)
        {
            constexpr int sw = (J + AHEAD + BACK) % WS;
            constexpr int ss = (J + AHEAD) % P;
            const i64 ra = q + AHEAD;
            if (ra < n1) {
#pragma unroll
                for (int v = 0; v < VEC; ++v) win[sw][v] = (WT)stg[ss][v];
                // 2. refill that staging slot with row q+AHEAD+P
                const i64 rn = ra + P;
                if (rn < n1) load_row(rn, stg[ss]);
            } else {
                const int h = (int)(ra - n1);  // < AHEAD: the line's first rows, saved before any write
#pragma unroll
                for (int hh = 0; hh < AHEAD; ++hh)
                    if (h == hh) {
#pragma unroll
                        for (int v = 0; v < VEC; ++v) win[sw][v] = (WT)head[hh][lane * VEC + v];
                    }
            }
        }
        // 3. FP64 foot location of row q, identical in every lane
        double d_prev1;
        const int leftNode = adv_foot(p, (int)i0, (int)q, d_prev1);
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, wave_uniform(d_prev1));
        if (bits != cached_bits) {
            double c[6];
            adv_lag_basis(__builtin_bit_cast(double, bits), c);
#pragma unroll
            for (int k = 0; k < 6; ++k) coef[k] = wave_uniform(c[k]);
            cached_bits = bits;
        }
        int d = wave_uniform((int)q - leftNode);
        if (d > half) d -= (int)n1;
        if (d < -half) d += (int)n1;
        // 4. stencil rows q-d-2 .. q-d+3 from the window, 5. store row q
        T out[VEC];
        interp_shift<T, WT, VEC, WS, J, BACK, S::DLO, S::DHI>(d, win, coef, out);
        if (lane_ok) M::store(slab + q * n2, out);
    };

    for (; base < n1; base += U) static_for<0, U>(row_step);
}

template <class T, class WT, int VEC, int SMAX, int P, int WPB, bool NT>
__global__ __launch_bounds__(64 * WPB) void adv_column_stream_kernel(T *__restrict__ data, StreamArgs a) {
    using S = StreamShape<SMAX, P>;
    __shared__ T head_lds[WPB][S::AHEAD][64 * VEC];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    // Persistent waves: unit = wave, wave + stride, ...  (all lanes of a wave
    // take the same trip count; there is no work-group barrier).
    for (i64 unit = (i64)blockIdx.x * WPB + wv; unit < a.total_units; unit += a.wave_stride)
        stream_unit<T, WT, VEC, SMAX, P, NT>(data, a, unit, lane, head_lds[wv]);
}

}  // namespace bkma
