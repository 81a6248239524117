// Column-pair advection kernel: the column stream of adv_stream.hpp for
// slabs whose backward shifts d = i1 - leftNode take at most two adjacent
// values {D, D+1} along the whole line.
//
// With a constant velocity per i0 slab the foot of row i1 lies s = v*dt/dx
// cells back, so leftNode = floor(i1 - s + e) with e a rounding error far
// below one cell: d = ceil(s - e) is the same for every row unless s is
// within e of an integer, and then it takes two adjacent values.  A plan
// scans every (i0, i1) once with the device's own foot arithmetic
// (slab_shift_scan_kernel) and keeps the smallest and largest d of each
// slab.  A slab with one shift D (almost every slab) runs the reference's
// six taps over rows i1-D-2 .. i1-D+3 with six-row windows; a slab with two
// shifts runs the seven-tap form below.
//
// The six taps of a row with shift d read rows i1-d-2 .. i1-d+3.  Over the
// seven rows i1-D-3 .. i1-D+3 they are the seven coefficients
//   d = D+1: (c0, c1, c2, c3, c4, c5, 0)
//   d = D  : (0, c0, c1, c2, c3, c4, c5)
// The extra product is an exact zero, so the separately rounded sum (FP64
// storage) and the fused sum (FP32 storage) keep the reference's value bit
// for bit on finite data, and the inner loop has no branch on the shift.
// One instance of the stream loop per D (selected once per work unit) fixes
// the stage lag (3 - D rows) and the window (seven rows per stage) at
// compile time.  Periodic wrap, in-place order, the head rows kept in LDS,
// the foot table and the FP64 intermediate stages follow adv_stream.hpp.
//
// Coefficient broadcast.  Each loop trip evaluates the foot and lag_basis
// of its rows, one row per lane.  DPP = false: every stage reads its row's
// seven coefficients back with v_readlane (14 VALU instructions per row and
// stage).  DPP = true (FP32 storage): the trip's table is copied into sets
// of 16 rows, each replicated in the four 16-lane rows of the wavefront
// (ds_bpermute), and the fused multiply-adds read the coefficient of table
// lane b straight from that set with the DPP modifier row_newbcast:b, so the
// broadcast costs no instruction of its own.  (For FP64 storage, whose
// separately rounded multiply has no DPP form, a 64-bit DPP move per
// coefficient measured no faster than v_readlane.)
#pragma once

#include "adv_stream.hpp"

namespace bkma {

constexpr int kPairTaps = 7;

// Stage windows of K taps (K = 6: one shift D; K = 7: shifts {D, D+1}).
// PV is the variant's prefetch depth; the six-tap instance rounds it down to
// a multiple of six so that one loop trip stays one window long.
// BACK may be negative (taps entirely ahead of the row, D < 4 - K): then no
// row before row 0 is needed and R = 0.  AHEAD may be 0 (D = 3): stage t
// then reads the row stage t-1 produced at the same position.
template <int K, int D, int PV, int NS> struct PairShape {
    static constexpr int BACK = D + K - 4;  // rows behind the output row read by its taps
    static constexpr int AHEAD = 3 - D;     // rows in front of it (stage lag)
    static constexpr int W = K;             // rows per stage window
    static constexpr int P = K == 7 ? PV : (PV >= 12 ? 12 : 6);
    static constexpr int U = clcm(W, P);
    static constexpr int R = NS * (BACK > 0 ? BACK : 0);
    static constexpr int H = NS * AHEAD;
    static constexpr int TB = NS * AHEAD;
    static_assert(K == 6 || K == 7, "six or seven taps");
    static_assert(AHEAD >= 0 && BACK + AHEAD == K - 1, "shift outside the supported range");
    static_assert(U + TB <= 64, "foot table must fit one wavefront");
};

// K-tap sum of VEC cells from window slots S0..S0+K-1 (mod K), in FP64:
// separately rounded multiply and add for FP64 storage, fused for FP32.
template <class T, class WT, class OT, int K, int VEC, int S0>
__device__ __forceinline__ void pair_taps(const WT (&win)[K][VEC], const double (&c)[kPairTaps], OT (&out)[VEC]) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        double acc = 0.;
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int k = 0; k < K; ++k) acc = acc + c[k] * (double)win[(S0 + k) % K][v];
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) acc = fma(c[k], (double)win[(S0 + k) % K][v], acc);
        }
        out[v] = (OT)acc;
    }
}

// The fused sum (FP32 storage) with the coefficients taken from lane B of
// each 16-lane row of `set`: v_fmac_f64 with the DPP source modifier
// row_newbcast:B.  The leading s_nop 1 covers the two wait states a DPP read
// needs after a VALU write of its source register.
template <class T, class WT, class OT, int K, int VEC, int S0, int B>
__device__ __forceinline__ void pair_taps_dpp(const WT (&win)[K][VEC], const double (&set)[kPairTaps],
                                              OT (&out)[VEC]) {
    static_assert(B >= 0 && B < 16, "row_newbcast lane");
    static_assert(sizeof(T) == 4, "DPP broadcast: fused (FP32 storage) sums only");
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        double w[kPairTaps];
#pragma unroll
        for (int k = 0; k < K; ++k) w[k] = (double)win[(S0 + k) % K][v];
        double acc = 0.;
        if constexpr (K == 6) {
            asm("s_nop 1\n\t"
                "v_fmac_f64_dpp %0, %1, %7 row_newbcast:%c13 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %0, %2, %8 row_newbcast:%c13 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %0, %3, %9 row_newbcast:%c13 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %0, %4, %10 row_newbcast:%c13 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %0, %5, %11 row_newbcast:%c13 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %0, %6, %12 row_newbcast:%c13 row_mask:0xf bank_mask:0xf"
                : "+v"(acc)
                : "v"(set[0]), "v"(set[1]), "v"(set[2]), "v"(set[3]), "v"(set[4]), "v"(set[5]), "v"(w[0]),
                  "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5]), "n"(B));
            out[v] = (OT)acc;
            continue;
        }
        asm("s_nop 1\n\t"
            "v_fmac_f64_dpp %0, %1, %8 row_newbcast:%c15 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %0, %2, %9 row_newbcast:%c15 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %0, %3, %10 row_newbcast:%c15 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %0, %4, %11 row_newbcast:%c15 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %0, %5, %12 row_newbcast:%c15 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %0, %6, %13 row_newbcast:%c15 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %0, %7, %14 row_newbcast:%c15 row_mask:0xf bank_mask:0xf"
            : "+v"(acc)
            : "v"(set[0]), "v"(set[1]), "v"(set[2]), "v"(set[3]), "v"(set[4]), "v"(set[5]), "v"(set[6]),
              "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5]), "v"(w[6]), "n"(B));
        out[v] = (OT)acc;
    }
}

// Value of `x` in lane `src` of the wavefront (ds_bpermute, 64-bit).
__device__ __forceinline__ double lane_gather(double x, int src) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(src * 4, (int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(src * 4, (int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The seven coefficients of a row with shift d, for the pair {D, D+1}.
template <int D> __device__ __forceinline__ void pair_coefs(int d, const double (&c)[6], double (&out)[kPairTaps]) {
    const bool late = d == D;  // taps start one row later in the window
    out[0] = late ? 0. : c[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) out[k] = late ? c[k - 1] : c[k];
    out[6] = late ? c[5] : 0.;
}

// One wavefront advances column tile `tile` of slab i0 by NS steps; every
// row's shift is D (K = 6) or D or D+1 (K = 7).
template <class T, class WT, int VEC, int PV, int NS, bool NT, bool DPP, int K, int D>
__device__ __forceinline__ void pair_unit(T *__restrict__ data, const StreamArgs &a, i64 i0, i64 tile, int lane,
                                          T (*head)[64 * VEC]) {
    using S = PairShape<K, D, PV, NS>;
    constexpr int AHEAD = S::AHEAD, BACK = S::BACK, W = S::W, P = S::P, U = S::U, R = S::R, H = S::H, TB = S::TB;
    constexpr int NSET = DPP ? (U + TB - AHEAD + 15) / 16 : 1;  // 16-row sets of the table the stages read
    constexpr int NM = NS > 1 ? NS - 1 : 1;
    using M = VecMem<T, VEC, NT>;
    const AdvDev &p = a.p;
    const i64 n1 = p.n1, n2 = p.n2;
    const i64 col = (tile * 64 + lane) * VEC;
    const bool lane_ok = col < n2;  // VEC divides n2: the whole vector is in range
    const bool full_tile = wave_uniform((int)((tile * 64 + 64) * VEC <= n2)) != 0;
    T *slab = data + (i0 * n1) * n2 + (lane_ok ? col : 0);
    const i64 first_end = R + n1;  // end of the positions read from HBM
    const i64 total = first_end + H;

    WT w0[W][VEC];
    double wm[NM][W][VEC];
    T stg[P][VEC];
#pragma unroll
    for (int s = 0; s < W; ++s)
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
    // A steady trip: every lane in range, no position of the trip in the
    // first or last trips' special cases (rows read before row 0, head rows,
    // outputs before the first stored row, positions past the line).
    auto load_row = [&](i64 pos, T (&dst)[VEC]) __attribute__((always_inline)) { M::load(slab + (pos - R) * n2, dst); };
    static_for<0, P>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        load_pos(s, stg[s]);
    });

    const int half = (int)(n1 / 2);
    double tab[kPairTaps];
    double tset[NSET][kPairTaps];

    i64 base = 0;
    auto trip = [&](auto steady_c) __attribute__((always_inline)) {
        // Foot table: lane l holds the K coefficients of the row at position
        // base - TB + l.
        {
            i64 row = base - TB + lane - R;
            if (row < 0) row += n1;
            if (row >= n1) row -= n1;
            if (row < 0 || row >= n1) row = 0;  // positions no stored value depends on
            double d_prev1, c[6];
            const int leftNode = adv_foot(p, (int)i0, (int)row, d_prev1);
            adv_lag_basis(d_prev1, c);
            if constexpr (K == 7) {
                int d = (int)row - leftNode;
                if (d > half) d -= (int)n1;
                if (d < -half) d += (int)n1;
                pair_coefs<D>(d, c, tab);
            } else {
#pragma unroll
                for (int k = 0; k < 6; ++k) tab[k] = c[k];
            }
        }
        if constexpr (DPP) {
            // set s, lane l: table entry 16*s + l%16
#pragma unroll
            for (int s = 0; s < NSET; ++s)
#pragma unroll
                for (int k = 0; k < K; ++k) tset[s][k] = lane_gather(tab[k], 16 * s + (lane & 15));
        }

        auto step = [&](auto jc, auto steady_c) __attribute__((always_inline)) {
            constexpr int J = decltype(jc)::value;
            constexpr bool STEADY = decltype(steady_c)::value;
            const i64 pin = base + J;
            if constexpr (!STEADY) {
                if (pin >= total) return;
            }
            // 1. position pin enters the stage-0 window
            {
                constexpr int sw = J % W, ss = J % P;
                if (STEADY || pin < first_end) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v) w0[sw][v] = (WT)stg[ss][v];
                    if constexpr (!STEADY) {
                        if (pin >= R && pin < R + H) {
#pragma unroll
                            for (int v = 0; v < VEC; ++v) head[pin - R][lane * VEC + v] = stg[ss][v];
                        }
                    }
                    // 2. refill the staging slot with position pin + P
                    if constexpr (STEADY)
                        load_row(pin + P, stg[ss]);
                    else if (pin + P < first_end)
                        load_pos(pin + P, stg[ss]);
                } else {
                    const int h = (int)(pin - first_end);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) w0[sw][v] = (WT)head[h][lane * VEC + v];
                }
            }
            // 3. stage t computes position pin - t*AHEAD from the K rows
            //    around it in stage t-1
            static_for<1, NS + 1>([&](auto tc) __attribute__((always_inline)) {
                constexpr int t = decltype(tc)::value;
                constexpr int OFF = J - t * AHEAD;
                constexpr int L = OFF + TB;
                constexpr int S0 = cmod(OFF - BACK, W);
                double c[kPairTaps];
                if constexpr (!DPP) {
#pragma unroll
                    for (int k = 0; k < K; ++k) c[k] = lane_value<L>(tab[k]);
                }
                auto taps = [&](const auto &win, auto &out) __attribute__((always_inline)) {
                    using WTT = std::remove_cv_t<std::remove_reference_t<decltype(win[0][0])>>;
                    using OTT = std::remove_reference_t<decltype(out[0])>;
                    if constexpr (DPP)
                        pair_taps_dpp<T, WTT, OTT, K, VEC, S0, L % 16>(win, tset[L / 16], out);
                    else
                        pair_taps<T, WTT, OTT, K, VEC, S0>(win, c, out);
                };
                if constexpr (t < NS) {
                    double o[VEC];
                    if constexpr (t == 1)
                        taps(w0, o);
                    else
                        taps(wm[t - 2], o);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) wm[t - 1][cmod(OFF, W)][v] = o[v];
                } else {
                    T o[VEC];
                    if constexpr (NS == 1)
                        taps(w0, o);
                    else
                        taps(wm[NS - 2], o);
                    const i64 pt = pin - NS * AHEAD;
                    if (STEADY || (pt >= R && pt < first_end && lane_ok)) M::store(slab + (pt - R) * n2, o);
                }
            });
        };
        static_for<0, U>([&](auto jc) __attribute__((always_inline)) { step(jc, steady_c); });
    };
    // Trips are taken in order; the steady ones (bases s_lo .. s_hi, all
    // multiples of U) run a body without the range checks, in a loop of its
    // own (one loop with both bodies needs far more registers).
    const i64 s_lo = full_tile ? (R + H + U - 1) / U * U : total;
    const i64 s_hi = full_tile ? first_end - U - P : -1;
    for (; base < total && base < s_lo; base += U) trip(std::false_type{});
    for (; base <= s_hi; base += U) trip(std::true_type{});
    for (; base < total; base += U) trip(std::false_type{});
}

// Selects the K-tap instance of D (wave-uniform) in [DLO, DHI].
template <class T, class WT, int VEC, int P, int NS, bool NT, bool DPP, int K, int DLO, int DHI>
__device__ __forceinline__ void pair_dispatch(int d, T *__restrict__ data, const StreamArgs &a, i64 i0, i64 tile,
                                              int lane, T (*head)[64 * VEC]) {
    if constexpr (DLO <= DHI) {
        if (d == DLO) {
            pair_unit<T, WT, VEC, P, NS, NT, DPP, K, DLO>(data, a, i0, tile, lane, head);
            return;
        }
        pair_dispatch<T, WT, VEC, P, NS, NT, DPP, K, DLO + 1, DHI>(d, data, a, i0, tile, lane, head);
    }
}

// Slabs with shifts in [-SMAX, SMAX] (SMAX <= 3): one shift D in
// [-SMAX, SMAX] runs the six-tap instance, two shifts {D, D+1} the seven-tap
// instance of D in [-SMAX, SMAX-1].  The head rows live in dynamic LDS,
// NS * a.pair_ahead rows per wave (pair_ahead = 3 - the smallest shift of the
// problem, the largest stage lag of its slabs).
template <class T, class WT, int VEC, int SMAX, int P, int NS, int WPB, bool NT, bool DPP>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(2))) void adv_column_pair_kernel(T *__restrict__ data, StreamArgs a) {
    static_assert(SMAX <= 3, "stage lag 3 - D must stay non-negative");
    extern __shared__ __align__(16) unsigned char pair_lds[];
    const int lane = threadIdx.x & 63;
    const int wv = wave_uniform((int)(threadIdx.x >> 6));  // unit indices and addresses stay scalar
    T(*head)[64 * VEC] = reinterpret_cast<T(*)[64 * VEC]>(pair_lds) + (i64)wv * NS * a.pair_ahead;
    for (i64 unit = (i64)blockIdx.x * WPB + wv; unit < a.total_units; unit += a.wave_stride) {
        const i64 i0 = unit / a.tiles_per_slab;
        const i64 tile = unit - i0 * a.tiles_per_slab;
        const int dlo = wave_uniform(a.slab_shift[2 * i0]);
        const int dhi = wave_uniform(a.slab_shift[2 * i0 + 1]);
        if (dlo == dhi)
            pair_dispatch<T, WT, VEC, P, NS, NT, DPP, 6, -SMAX, SMAX>(dlo, data, a, i0, tile, lane, head);
        else
            pair_dispatch<T, WT, VEC, P, NS, NT, DPP, 7, -SMAX, SMAX - 1>(dlo, data, a, i0, tile, lane, head);
    }
}

// Dynamic LDS of one work-group of a pair variant: the head rows of its waves.
template <class T, int VEC, int WPB> inline size_t pair_lds_bytes(int ns, int ahead) {
    return (size_t)WPB * ns * (ahead > 0 ? ahead : 0) * 64 * VEC * sizeof(T) + 16;
}

// Per-slab shift range with the kernels' own foot arithmetic: shift[2*i0]
// = min d, shift[2*i0+1] = max d, and stats = {min over slabs of min d, max
// over slabs of max d, max over slabs of (max d - min d)} (stats preset to
// {INT_MAX, INT_MIN, 0}).
__global__ void slab_shift_scan_kernel(AdvDev p, int *__restrict__ shift, int *__restrict__ stats);


// ---------------------------------------------------------------------------
// Host side: launchers for steps-per-pass 1 .. NSMAX of one variant
// ---------------------------------------------------------------------------
template <class T, class WT, int VEC, int SMAX, int P, int WPB, bool NT, bool DPP, int NS>
void launch_pair(void *data, const StreamArgs &a, unsigned grid, hipStream_t s, int ns) {
    if constexpr (NS >= 1) {
        if (ns == NS) {
            const size_t lds = pair_lds_bytes<T, VEC, WPB>(NS, a.pair_ahead);
            hipLaunchKernelGGL((adv_column_pair_kernel<T, WT, VEC, SMAX, P, NS, WPB, NT, DPP>), dim3(grid),
                               dim3(64 * WPB), lds, s, reinterpret_cast<T *>(data), a);
            return;
        }
        launch_pair<T, WT, VEC, SMAX, P, WPB, NT, DPP, NS - 1>(data, a, grid, s, ns);
    }
}

template <class T, class WT, int VEC, int SMAX, int P, int WPB, bool NT, bool DPP, int NS>
int occupancy_pair(int ns, int ahead) {
    if constexpr (NS >= 1) {
        if (ns == NS) {
            int blocks = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &blocks, adv_column_pair_kernel<T, WT, VEC, SMAX, P, NS, WPB, NT, DPP>, 64 * WPB,
                    pair_lds_bytes<T, VEC, WPB>(NS, ahead)) != hipSuccess)
                return 0;
            return blocks;
        }
        return occupancy_pair<T, WT, VEC, SMAX, P, WPB, NT, DPP, NS - 1>(ns, ahead);
    }
    return 0;
}

#define BKMA_PAIR_VARIANT(T, WT, DT, V, S, P, W, NT, DPP, NSMAX)                                                \
    StreamVariant {                                                                                            \
        DT, 1, V, S, P, W, NT, sizeof(WT) > sizeof(T), NSMAX, &launch_pair<T, WT, V, S, P, W, NT, DPP, NSMAX>, \
            &occupancy_pair<T, WT, V, S, P, W, NT, DPP, NSMAX>, DPP                                            \
    }

}  // namespace bkma
