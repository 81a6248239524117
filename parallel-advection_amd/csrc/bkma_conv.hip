// Conv1d (textbook semantics, torch.nn.functional.conv1d with stride 1 and
// no padding) written in place over its input:
//   y[b][oc][j] = bias[oc] + sum_ic sum_k w[oc][ic][k] * x[b][ic][j + k],
//   j in [0, L-K+1) overwrites x[b][oc][j]; x[b][oc][L-K+1 .. L) is kept.
//
// In-place safety: y[b][:, J] overwrites x[b][:, J], which every output
// channel still needs.  Every kernel walks columns in increasing tiles
// [j0, j0+TN): a tile reads columns [j0, j0+TN+K-1) and writes [j0, j0+TN),
// so reads run ahead of writes.  The direct kernels give one work-group one
// batch element; the MFMA kernels give one work-group a contiguous chunk of
// the flattened (batch, tile) sequence, and the K-1 columns past a chunk's
// last tile are snapshotted first (conv_halo_kernel).
//
// Kernels
//  - conv_direct_kernel<T>: FP64 / FP32 FMA reference kernels (FP64 keeps the
//    oracle's order: bias, then ic outer, k inner, separate multiply/add).
//  - conv_mfma_kernel<TIO, NPROD>: implicit GEMM on v_mfma_f32_32x32x16_bf16.
//    GEMM: Y[oc][j] = sum_r W[oc][r] X[r][j] with r = k*C + ic, so 8
//    consecutive r are 8 consecutive channels at one tap and one ds_read_b128
//    of the channel-last LDS tile Xs[jj][ic] gives a B fragment.  Each wave
//    owns 32 output channels and keeps its W fragments in registers for the
//    whole kernel (weight stationary).  NPROD = 1: bf16 x bf16; NPROD = 3:
//    split precision (x = xh + xl, w = wh + wl, y += xh wh + xh wl + xl wh).
//  - conv_tap_bf16_kernel: the BF16 mode with BF16 storage.  One B fragment
//    of X feeds the K taps' MFMAs (K accumulators over the same columns,
//    combined by DPP lane shifts), so it reads K times fewer LDS fragments
//    than conv_mfma_kernel; its loads and stores are whole 64-byte pieces.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "mfma_ops.hpp"

namespace bkma {


#ifndef BKMA_CONV_PAD
#define BKMA_CONV_PAD 8
#endif
constexpr int kConvPad = BKMA_CONV_PAD;  // bf16 elements of padding per LDS row (keeps 16-byte alignment)

template <class T> __device__ __forceinline__ float io_to_f32(T v) { return (float)v; }
template <> __device__ __forceinline__ float io_to_f32<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <class T> __device__ __forceinline__ T f32_to_io(float v) { return (T)v; }
template <> __device__ __forceinline__ __hip_bfloat16 f32_to_io<__hip_bfloat16>(float v) {
    return __builtin_bit_cast(__hip_bfloat16, (__bf16)v);  // round to nearest even (v_cvt_pk_bf16_f32)
}

// ---------------------------------------------------------------------------
// Direct kernels (FP64 / FP32)
// ---------------------------------------------------------------------------
constexpr int kDirTN = 64;   // columns per tile
constexpr int kDirOC = 4;    // output channels per thread
constexpr int kDirJ = 8;     // columns per thread

template <class T>
__global__ __launch_bounds__(256) void conv_direct_kernel(T *__restrict__ x, const T *__restrict__ w,
                                                          const T *__restrict__ bias, int C, int L, int K) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *xs = reinterpret_cast<T *>(lds_raw);  // [C][JW]
    const int JW = kDirTN + K - 1;
    const int Lout = L - K + 1;
    T *xb = x + (i64)blockIdx.x * C * L;
    const int t = threadIdx.x;
    const int jb = t % (kDirTN / kDirJ);                 // 8 column blocks
    const int ob = t / (kDirTN / kDirJ);                 // 32 channel blocks of 4
    const int oc_groups = (C + kDirOC * 32 - 1) / (kDirOC * 32);
    for (int j0 = 0; j0 < Lout; j0 += kDirTN) {
        __syncthreads();
        for (int e = t; e < C * JW; e += 256) {
            const int ic = e / JW, jj = e - ic * JW;
            const int j = j0 + jj;
            xs[e] = j < L ? xb[(i64)ic * L + j] : (T)0;
        }
        __syncthreads();
        for (int g = 0; g < oc_groups; ++g) {
            const int oc0 = (g * 32 + ob) * kDirOC;
            T acc[kDirOC][kDirJ];
#pragma unroll
            for (int a = 0; a < kDirOC; ++a)
#pragma unroll
                for (int c = 0; c < kDirJ; ++c) acc[a][c] = oc0 + a < C ? bias[oc0 + a] : (T)0;
            for (int ic = 0; ic < C; ++ic) {
                const T *xr = xs + ic * JW + jb * kDirJ;
                for (int k = 0; k < K; ++k) {
                    T wv[kDirOC];
#pragma unroll
                    for (int a = 0; a < kDirOC; ++a) wv[a] = oc0 + a < C ? w[((i64)(oc0 + a) * C + ic) * K + k] : (T)0;
#pragma unroll
                    for (int c = 0; c < kDirJ; ++c) {
                        const T xv = xr[c + k];
#pragma unroll
                        for (int a = 0; a < kDirOC; ++a) acc[a][c] = acc[a][c] + wv[a] * xv;
                    }
                }
            }
            // Stores land in columns [j0, j0+TN), which no later tile reads.
#pragma unroll
            for (int a = 0; a < kDirOC; ++a)
#pragma unroll
                for (int c = 0; c < kDirJ; ++c) {
                    const int j = j0 + jb * kDirJ + c;
                    if (oc0 + a < C && j < Lout) xb[(i64)(oc0 + a) * L + j] = acc[a][c];
                }
        }
    }
}

// ---------------------------------------------------------------------------
// MFMA kernel
// ---------------------------------------------------------------------------
// W fragments, packed once per plan: for wave w (channels 32w..32w+31),
// k-step s (r in [16s, 16s+16)) and lane l, 8 bf16 values
//   A[oc = 32w + l%32][r = 16s + 8*(l/32) + e], e = 0..7,  r = k*C + ic.
__global__ void conv_pack_weights_kernel(const float *__restrict__ w, int C, int K, __bf16 *__restrict__ hi,
                                         __bf16 *__restrict__ lo) {
    const int KS = C * K / 16;
    const i64 total = (i64)(C / 32) * KS * 64 * 8;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const int e = (int)(idx % 8);
        const int l = (int)((idx / 8) % 64);
        const int s = (int)((idx / 512) % KS);
        const int wv = (int)(idx / (512 * (i64)KS));
        const int oc = 32 * wv + l % 32;
        const int r = 16 * s + 8 * (l / 32) + e;
        const int k = r / C, ic = r - k * C;
        const float v = w[((i64)oc * C + ic) * K + k];
        const __bf16 h = to_bf16(v);
        hi[idx] = h;
        if (lo) lo[idx] = to_bf16(v - (float)h);
    }
}

// In place: a work-group takes a contiguous chunk of tiles (flattened over
// (b, t)) in increasing order, so a tile's K-1 columns beyond its own -- the
// first columns of the next tile -- are read before that tile is written,
// except for the last tile of a chunk, whose neighbour belongs to another
// work-group.  Those K-1 columns are copied first into halo[chunk][ic][K-1].
template <class TIO>
__global__ void conv_halo_kernel(const TIO *__restrict__ x, TIO *__restrict__ halo, int C, int L, int K, int TN,
                                 int tiles, int total_tiles, int chunk) {
    const int c = blockIdx.x;
    const int last = min((c + 1) * chunk, total_tiles) - 1;
    if (last < c * chunk) return;  // empty chunk
    const int b = last / tiles, t = last - b * tiles;
    if (t >= tiles - 1) return;  // the row's last tile reads no neighbour
    const int H = K - 1;
    for (int e = threadIdx.x; e < C * H; e += blockDim.x) {
        const int ic = e / H, kk = e - ic * H;
        halo[(i64)c * C * H + e] = x[((i64)b * C + ic) * L + (i64)(t + 1) * TN + kk];
    }
}

// Element c of a vector load as the io type.  The element is copied out
// first: a bit_cast applied to the subscript expression itself read element
// 0 for every c with this compiler.
template <class TIO, class V> __device__ __forceinline__ TIO io_elem(const V &v, int c) {
    const auto e = v[c];
    return __builtin_bit_cast(TIO, e);
}

template <class T, int N> struct IoVec;
template <> struct IoVec<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct IoVec<__hip_bfloat16, 4> { typedef unsigned short type __attribute__((ext_vector_type(4))); };

// Columns j .. j+3 of a row of length L (j a multiple of 4), clamped into
// the row and unmasked: the caller zeroes columns >= L when it consumes them,
// so the loads stay in flight until then.  load_quad_vec keeps the quad as
// one packed vector (2 VGPRs for bf16).
template <class TIO, bool VL>
__device__ __forceinline__ void load_quad_vec(const TIO *row, int j, int L, typename IoVec<TIO, 4>::type &out) {
    using V4 = typename IoVec<TIO, 4>::type;
    if constexpr (VL) {
        out = *reinterpret_cast<const V4 *>(row + (j < L ? j : L - 4));
    } else {
        using E = typename std::conditional<sizeof(TIO) == 4, float, unsigned short>::type;
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = __builtin_bit_cast(E, row[j + c < L ? j + c : L - 1]);
    }
}

template <class TIO, bool VL> __device__ __forceinline__ void load_quad_raw(const TIO *row, int j, int L, TIO (&out)[4]) {
    if constexpr (VL) {
        using V4 = typename IoVec<TIO, 4>::type;
        const V4 v = *reinterpret_cast<const V4 *>(row + (j < L ? j : L - 4));
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = io_elem<TIO>(v, c);
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = row[j + c < L ? j + c : L - 1];
    }
}

// Persistent implicit-GEMM kernel.  A work-group (C/32 waves) owns all C
// output channels of a tile of TN columns of one batch element; wave w keeps
// the W fragments of channels 32w..32w+31 in registers.  Tiles are taken in
// a grid-stride loop, double-buffered in LDS: the next tile's global loads
// are issued before the current tile's MFMAs and written to the other LDS
// buffer after them, one barrier per tile.
// LDS tile: Xs[jj][ic] bf16, row stride C + kConvPad (padding against bank
// conflicts of the B-fragment reads).
template <class TIO, int NPROD, int C, int K, int TN, bool VL>
__global__ __launch_bounds__(2 * C, 1) void conv_mfma_kernel(TIO *__restrict__ x, const TIO *__restrict__ halo,
                                                           const __bf16 *__restrict__ whi,
                                                           const __bf16 *__restrict__ wlo,
                                                           const float *__restrict__ bias, int L, int tiles,
                                                           int total_tiles, int chunk) {
    constexpr int NW = C / 32;
    constexpr int NT = 64 * NW;
    constexpr int KS = C * K / 16;
    constexpr bool XLO = NPROD == 3 && !std::is_same<TIO, __hip_bfloat16>::value;  // bf16 input: x = xh exactly
    constexpr int JW = TN + K - 1;
    constexpr int RS = C + kConvPad;
    constexpr int NPART = XLO ? 2 : 1;
    constexpr int BUF = JW * RS;  // bf16 elements of one part of one buffer
    constexpr int NSUB = TN / 32;
    constexpr int NQ = TN / 4;                // 4-column groups
    constexpr int NTASK = (C / 8) * NQ;       // (8-channel group, 4-column group) staging tasks
    static_assert(NTASK % NT == 0, "staging tasks must divide evenly");
    constexpr int TPT = NTASK / NT;
    constexpr int NH = C * (K - 1);
    static_assert(NH % NT == 0, "halo elements must divide evenly");
    constexpr int HPT = NH / NT;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16 *xs = reinterpret_cast<__bf16 *>(lds_raw);  // [2][NPART][JW][RS]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int Lout = L - K + 1;

    int tile = blockIdx.x * chunk;
    const int tile_end = min(tile + chunk, total_tiles);
    if (tile >= tile_end) return;  // empty chunk (uniform over the work-group)

    bf16x8_t wa[KS];
    bf16x8_t wb[NPROD == 3 ? KS : 1];
    {
        const bf16x8_t *src = reinterpret_cast<const bf16x8_t *>(whi) + ((i64)wv * KS) * 64 + lane;
#pragma unroll
        for (int s = 0; s < KS; ++s) wa[s] = src[s * 64];
        if constexpr (NPROD == 3) {
            const bf16x8_t *srl = reinterpret_cast<const bf16x8_t *>(wlo) + ((i64)wv * KS) * 64 + lane;
#pragma unroll
            for (int s = 0; s < KS; ++s) wb[s] = srl[s * 64];
        }
    }

    float bias_r[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) bias_r[r] = bias[32 * wv + 8 * (r / 4) + 4 * (lane / 32) + (r % 4)];

    typename IoVec<TIO, 4>::type pf[TPT][8];  // raw quads of the next tile
    TIO ph[HPT];
    auto load_tile = [&](int tl) __attribute__((always_inline)) {
        const int b = tl / tiles, t = tl - b * tiles;
        const int j0 = t * TN;
        const TIO *xb = x + (i64)b * C * L;
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            const int task = tid + u * NT;
            const int q = task % NQ, g = task / NQ;
            const int j = j0 + 4 * q;
#pragma unroll
            for (int i = 0; i < 8; ++i) load_quad_vec<TIO, VL>(xb + (i64)(8 * g + i) * L, j, L, pf[u][i]);
        }
        const bool snap = tl == tile_end - 1 && t < tiles - 1;
#pragma unroll
        for (int u = 0; u < HPT; ++u) {
            const int e = tid + u * NT;
            const int ic = e / (K - 1), kk = e - ic * (K - 1);
            const int j = j0 + TN + kk;
            ph[u] = *(snap ? halo + (i64)blockIdx.x * NH + e : xb + (i64)ic * L + (j < L ? j : L - 1));
        }
    };
    // Converts the loaded tile tl (zero past L) into LDS buffer buf.
    auto stage_tile = [&](int buf, int tl) __attribute__((always_inline)) {
        __bf16 *hi = xs + (buf * NPART) * BUF;
        __bf16 *lo = hi + BUF;
        const int t = tl % tiles;
        const int j0 = t * TN;
        const bool snap = tl == tile_end - 1 && t < tiles - 1;
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            const int task = tid + u * NT;
            const int q = task % NQ, g = task / NQ;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool in = j0 + 4 * q + c < L;
                bf16x8_t hv, lv;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if constexpr (std::is_same<TIO, __hip_bfloat16>::value) {
                        hv[i] = in ? __builtin_bit_cast(__bf16, io_elem<TIO>(pf[u][i], c)) : (__bf16)0.0f;
                    } else {
                        const float f = in ? io_to_f32(io_elem<TIO>(pf[u][i], c)) : 0.0f;
                        const __bf16 h = to_bf16(f);
                        hv[i] = h;
                        if constexpr (XLO) lv[i] = to_bf16(f - (float)h);
                    }
                }
                const int off = (4 * q + c) * RS + 8 * g;
                *reinterpret_cast<bf16x8_t *>(hi + off) = hv;
                if constexpr (XLO) *reinterpret_cast<bf16x8_t *>(lo + off) = lv;
            }
        }
#pragma unroll
        for (int u = 0; u < HPT; ++u) {
            const int e = tid + u * NT;
            const int ic = e / (K - 1), kk = e - ic * (K - 1);
            const float f = snap || j0 + TN + kk < L ? io_to_f32(ph[u]) : 0.0f;
            const __bf16 h = to_bf16(f);
            hi[(TN + kk) * RS + ic] = h;
            if constexpr (XLO) lo[(TN + kk) * RS + ic] = to_bf16(f - (float)h);
        }
    };

    load_tile(tile);
    stage_tile(0, tile);
    __syncthreads();
    const int jl = lane % 32, kh = 8 * (lane / 32);
    for (int it = 0; tile < tile_end; ++it) {
        const int next = tile + 1;
        if (next < tile_end) load_tile(next);
        const __bf16 *hi = xs + ((it & 1) * NPART) * BUF;
        const __bf16 *lo = hi + BUF;
        f32x16_t acc[NSUB];
#pragma unroll
        for (int n = 0; n < NSUB; ++n)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[n][r] = 0.0f;
        // B fragments are read PD steps ahead of their MFMA (a ring of PD+1),
        // so LDS latency overlaps the MFMAs of the previous steps.
        constexpr int NSTEP = KS * NSUB, PD = 2;
        bf16x8_t bq[PD + 1], lq[XLO ? PD + 1 : 1];
        auto bread = [&](auto ic_) __attribute__((always_inline)) {
            constexpr int idx = decltype(ic_)::value;
            constexpr int s = idx / NSUB, n = idx % NSUB;
            constexpr int k = (16 * s) / C, icb = 16 * s - k * C;
            const int off = (32 * n + jl + k) * RS + icb + kh;
            bq[idx % (PD + 1)] = *reinterpret_cast<const bf16x8_t *>(hi + off);
            if constexpr (XLO) lq[idx % (PD + 1)] = *reinterpret_cast<const bf16x8_t *>(lo + off);
        };
        static_for<0, PD>(bread);
        static_for<0, NSTEP>([&](auto ic_) __attribute__((always_inline)) {
            constexpr int idx = decltype(ic_)::value;
            constexpr int s = idx / NSUB, n = idx % NSUB;
            if constexpr (idx + PD < NSTEP) bread(std::integral_constant<int, idx + PD>{});
            const bf16x8_t bh = bq[idx % (PD + 1)];
            acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], bh, acc[n], 0, 0, 0);
            if constexpr (XLO)
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], lq[idx % (PD + 1)], acc[n], 0, 0, 0);
            if constexpr (NPROD == 3) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wb[s], bh, acc[n], 0, 0, 0);
        });
        // D[m][n]: lane l, register r -> m = 8*(r/4) + 4*(l/32) + r%4, n = l%32.
        {
            const int b = tile / tiles, t = tile - b * tiles;
            TIO *xb = x + (i64)b * C * L;
#pragma unroll
            for (int n = 0; n < NSUB; ++n) {
                const int j = t * TN + 32 * n + jl;
                if (j < Lout) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int oc = 32 * wv + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
                        xb[(i64)oc * L + j] = f32_to_io<TIO>(acc[n][r] + bias_r[r]);
                    }
                }
            }
        }
        if (next < tile_end) stage_tile((it + 1) & 1, next);
        __syncthreads();
        tile = next;
    }
}

// Staging task u of conv_tap_bf16_kernel loads in block 2u (at most TB-3)
// and is written to LDS two blocks later.
template <int TB> constexpr int tap_ld_blk(int u) { return 2 * u < TB - 3 ? 2 * u : TB - 3; }

// Tap-reuse implicit GEMM for BF16 storage on v_mfma_f32_32x32x16_bf16 (FP32
// accumulation), for shapes whose weight fragments stay in registers
// (K*C/4 per lane).  Per tap k, D_k = W_k X with A = W_k (the wave's 32
// output channels x 16 input channels, packed as in conv_mfma_kernel and kept
// in AGPRs for the whole kernel) and B = X (16 input channels x 32 columns,
// one ds_read_b128 of the channel-last LDS tile).  The B fragment of a k-step
// is shared by the K taps' MFMAs, which accumulate into K accumulators over
// the same 32 unshifted columns: a block of 32 columns reads C/16 fragments
// for K*C/16 MFMAs.  Lane l of every D_k holds column c0 + l%32.  The output
//   y[oc][c0 + j] = bias[oc] + sum_k D_k[oc][c0 + j + k],   j < OB = 33 - K,
// is combined in place by Horner's rule, y = D_0 + S(D_1 + S(D_2 + ...)),
// where S moves every column down by one lane: one DPP add (wave_shl:1, lane
// l reads lane l+1) per register and level.  Lanes 31 and 63 read values of
// another column set, which only reach the columns j >= OB that are not
// stored (consecutive blocks overlap by K - 1 columns), so no partial sum
// crosses a block.  The bias is added before the conversion to bf16.
//
// Schedule.  With one wave per SIMD the wave overlaps its own VALU and memory
// work with its MFMAs (an MFMA occupies the matrix core for 32 cycles, in
// which the wave can issue about six independent VALU instructions).  The
// blocks of a tile are unrolled; after each MFMA of block b comes one slice
// of the other work: block b-1's combination and stores (the other
// accumulator set), the B-fragment read two k-steps ahead, and the staging
// of the next tile into the other LDS buffer.  Scheduling barriers keep that
// order.  Block b-1's accumulators are read by VALU instructions only after
// an s_nop fence of 20 wait states that follows block b's first MFMA.  The
// last block's combination runs in the next tile's first block.
//
// Memory.  Global loads and stores are buffer instructions with one
// descriptor per batch element.  Tiles start at multiples of 64 bytes of a
// row, so the staging loads read whole 64-byte pieces, and the outputs go
// through a per-wave LDS ring so that they are stored as whole 64-byte
// pieces too (fewer, full requests to the L2).  A store that must not be
// written (past L-K+1) gets an offset beyond the descriptor's range, which
// the hardware drops: the block code has no branches.
//
// LDS.  A row holds C bf16 and 32 bytes of padding (18 chunks of 16 bytes for
// C = 128, so row r starts at bank group 2r mod 8), and its channel chunks
// are XOR-swizzled by row bits 2..5.  A B-fragment read (8 lanes: 8
// consecutive rows, one chunk) and a staging write (8 lanes: rows 4q+c over 8
// consecutive q, one chunk) then both hit 8 different groups of 4 banks.
// Requirements (checked by the host): L % 8 == 0 and C*L*2 < 2^31.
template <int C, int K, int TB>
__global__ __launch_bounds__(2 * C, 1) void conv_tap_bf16_kernel(__hip_bfloat16 *__restrict__ x,
                                                               const __hip_bfloat16 *__restrict__ halo,
                                                               const __bf16 *__restrict__ wq,
                                                               const float *__restrict__ bias, int L, int tiles,
                                                               int total_tiles, int chunk) {
    constexpr int NW = C / 32;
    constexpr int NT = 64 * NW;
    constexpr int KS = C / 16;       // k-steps (16 channels) per tap
    constexpr int OB = 33 - K;       // output columns per 32-column block
    constexpr int TO = TB * OB;      // output columns per tile
    constexpr int NXC = TO + K - 1;  // staged input columns per tile
    constexpr int RSB = 2 * C + 32;  // bytes per LDS row
    constexpr int BUFB = NXC * RSB;  // bytes per LDS buffer
    constexpr int NQ = NXC / 4;      // 4-column groups per tile
    constexpr int NTASK = (C / 8) * NQ;         // staging tasks (8 channels x 4 columns)
    constexpr int TPT = (NTASK + NT - 1) / NT;  // staging tasks per thread (the last may repeat one)
    static_assert(C % 128 == 0, "at least 16 chunks of 16 bytes per LDS row (the swizzle takes 4 bits)");
    static_assert(K >= 2 && (K - 1) % 4 == 0 && TO % 4 == 0, "halo and tile columns are whole quads");
    static_assert(TB % 2 == 0, "the two accumulator sets alternate over the blocks of a tile");
    static_assert(K + 3 <= KS, "k-steps: addressing, K-1 combination levels, stores, staging store and loads");
    static_assert(TPT >= 1 && (TPT < 2 || 2 * (TPT - 2) < TB - 3) && tap_ld_blk<TB>(TPT - 1) + 2 < TB,
                  "staging tasks load at blocks 0, 2, 4, ... (the last at TB-3 at most) and store two blocks later");

    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int n = lane & 31, h = lane >> 5;
    const int Lout = L - K + 1;
    const unsigned num_records = (unsigned)(C * L * 2);
    constexpr unsigned kDrop = 0x80000000u;  // store offset beyond every descriptor's range

    int tile = blockIdx.x * chunk;
    const int tile_end = min(tile + chunk, total_tiles);
    if (tile >= tile_end) return;  // empty chunk (uniform over the work-group)

    bf16x8_t wf[K * KS];  // tap k, k-step s at k * KS + s (the packed order r = k*C + ic)
    {
        const bf16x8_t *src = reinterpret_cast<const bf16x8_t *>(wq) + ((i64)wv * K * KS) * 64 + lane;
#pragma unroll
        for (int i = 0; i < K * KS; ++i) wf[i] = src[i * 64];
    }
    // D lane l, register r: output channel 32*wv + 8*(r/4) + 4*h + r%4.
    float bias_r[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) bias_r[r] = bias[32 * wv + 8 * (r / 4) + 4 * h + (r % 4)];

    // Staging: thread tid moves channels 8g .. 8g+7 (g = tid / 16) of the
    // 4-column groups q = tid % 16 + 16u, u < TPT (16 lanes read 128
    // contiguous bytes of a row).  Groups past the tile repeat its last one.
    // Row q of the swizzle is q % 16 for every u, so the LDS offset of task u
    // is st_lds + u * 64 * RSB.
    static_assert(NT == 256 && C / 8 == 16 && 16 * (TPT - 1) < NQ && NQ <= 16 * TPT,
                  "staging map: 16 channel groups x 16 column groups per pass");
    const int sg = tid >> 4, sq = tid & 15;
    const int st_row = 8 * sg * L * 2;
    const int st_lds = 4 * sq * RSB + 16 * (sg ^ sq);
    const int q_last = min(sq + 16 * (TPT - 1), NQ - 1);
    const int st_lds_last = 4 * q_last * RSB + 16 * (sg ^ (q_last & 15));
    // Output staging.  Tiles are whole 32-column sectors (64 bytes of a row),
    // and a wave stores each sector of its 32 channel rows once it is
    // complete, as 64-byte row pieces.  The combined block (lane n holding
    // column n of channels 8*(r/4) + 4*h + r%4) is written as bf16 to a
    // per-wave ring of 64 columns x 32 channel rows in LDS (after the two
    // input buffers, OUTRS bytes per row); block p completes sector p-1 of
    // its tile (28(p+1) >= 32p for p <= 7), which is read back with lane l
    // holding channel l/4 + 16i, columns 8*(l%4) .. +7 of the sector, and
    // stored with one 16-byte store per lane and i (an 8-byte store covers
    // the 4 columns before L-K+1 when a piece straddles it).
    static_assert(TO % 32 == 0 && TO / 32 == TB - 1 && OB >= 16, "a tile is TB-1 sectors; block p completes sector p-1");
    constexpr int OUTRS = 192;  // bytes per ring row: 128 + 64 (rows 1 apart in different bank halves)
    const int out_base = 2 * BUFB + wv * 32 * OUTRS;
    const int sc_wr = out_base + 4 * h * OUTRS;  // + 2 * ((c0 + n) & 63) + m * OUTRS
    const int sc_rd = out_base + (lane >> 2) * OUTRS + 16 * (lane & 3);  // + i*16*OUTRS + 2*(32s & 63)
    const unsigned st_lane = (unsigned)(((32 * wv + (lane >> 2)) * L + 8 * (lane & 3)) * 2);
    auto rsrc_of = [&](int tl) __attribute__((always_inline)) {
        const int b = tl / tiles;
        return __builtin_amdgcn_make_buffer_rsrc(x + (i64)b * C * L, 0, (int)num_records, 0x00020000);
    };
    // Task u of a tile at column tcol: 8 channels x 4 columns as 8 row quads.
    auto stage_load = [&](__amdgpu_buffer_rsrc_t rs, int tcol, int u, u32x2_t(&dst)[8]) __attribute__((always_inline)) {
        // columns past the row end are clamped into it: they only feed masked outputs
        const int col = u == TPT - 1 ? 4 * q_last : 4 * sq + 64 * u;
        const unsigned voff = (unsigned)(st_row + 2 * min(tcol + col, L - 4));
#pragma unroll
        for (int i = 0; i < 8; ++i) dst[i] = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, i * L * 2, 0);
    };
    // Column c of task u: dword d holds channels 2d (low half) and 2d+1.
    auto stage_store_col = [&](unsigned char *buf, int u, int c, const u32x2_t(&src)[8]) __attribute__((always_inline)) {
        u32x4_t v;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const unsigned lo = src[2 * d][c >> 1], hi = src[2 * d + 1][c >> 1];
            v[d] = ((c & 1) ? ((lo >> 16) | (hi & 0xFFFF0000u)) : ((lo & 0xFFFFu) | (hi << 16)));
        }
        *reinterpret_cast<u32x4_t *>(buf + (u == TPT - 1 ? st_lds_last : st_lds + u * 64 * RSB) + c * RSB) = v;
    };
    // Registers r, r+1 of the combined block at local column c0, bias added,
    // to the ring (channel rows 8*(r/4) + r%4 and the next one of this lane's
    // group).  Lanes n >= OB write columns of the next block, which that
    // block overwrites.
    auto ring_pair = [&](const f32x16_t &y, int r, int c0) __attribute__((always_inline)) {
        const unsigned p = pack_bf16x2(y[r] + bias_r[r], y[r + 1] + bias_r[r + 1]);
        const int m = 8 * (r / 4) + (r % 4);
        unsigned char *dst = lds_raw + sc_wr + 2 * ((c0 + n) & 63);
        *reinterpret_cast<unsigned short *>(dst + m * OUTRS) = (unsigned short)p;
        *reinterpret_cast<unsigned short *>(dst + (m + 1) * OUTRS) = (unsigned short)(p >> 16);
    };
    auto ring_read = [&](int sec, int i) __attribute__((always_inline)) {
        return *reinterpret_cast<const u32x4_t *>(lds_raw + sc_rd + i * 16 * OUTRS + 2 * ((32 * sec) & 63));
    };
    // Half i of sector sec of the tile at column tcol (lim = L-K+1 - tcol):
    // a lane's 8 columns when all lie below lim, else their first 4 when
    // those do; the other stores are dropped.
    auto store_sector = [&](const u32x4_t &v, int sec, int i, __amdgpu_buffer_rsrc_t rs, int tcol, int lim)
                            __attribute__((always_inline)) {
        const int c = 32 * sec + 8 * (lane & 3);
        const unsigned off = st_lane + (unsigned)((i * 16 * L + tcol + 32 * sec) * 2);
        __builtin_amdgcn_raw_buffer_store_b128(v, rs, (int)(c + 8 <= lim ? off : kDrop), 0, 0);
        const u32x2_t lo = {v[0], v[1]};
        __builtin_amdgcn_raw_buffer_store_b64(lo, rs, (int)(c + 8 > lim && c + 4 <= lim ? off : kDrop), 0, 0);
    };
    // y[kk] += S(y[kk + 1]) for registers r0 .. r0+NR-1 (DPP wave_shl:1).
    auto horner = [&](f32x16_t &yk, const f32x16_t &y1, int r0, int nr) __attribute__((always_inline)) {
#pragma unroll
        for (int r = r0; r < r0 + nr; ++r) {
            const float v = y1[r];
            const float sv = __builtin_bit_cast(
                float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
            yk[r] += sv;
        }
    };

    // Initial tile: staged synchronously into buffer 0.
    {
        const __amdgpu_buffer_rsrc_t rs = rsrc_of(tile);
        const int tcol = (tile % tiles) * TO;
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            u32x2_t ld[8];
            stage_load(rs, tcol, u, ld);
#pragma unroll
            for (int c = 0; c < 4; ++c) stage_store_col(lds_raw, u, c, ld);
        }
    }
    // Every load issued so far (weights, bias, the first tile) has arrived:
    // without this wait the loop would wait for them in every tile, and with
    // them for all the stores of the tile before.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();

    f32x16_t acc0[K], acc1[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc1[k][r] = 0.0f;
    // The combination of the previous block may belong to the previous tile:
    // its descriptor and column base.  Nothing is stored before the first block.
    __amdgpu_buffer_rsrc_t rs_prev = rsrc_of(tile);
    int tcol_prev = 0, lim_prev = INT_MIN;

    for (int it = 0; tile < tile_end; ++it) {
        const int t = tile % tiles;
        unsigned char *const cur = lds_raw + (it & 1) * BUFB;
        unsigned char *const nxt = lds_raw + ((it + 1) & 1) * BUFB;
        if (tile == tile_end - 1 && t < tiles - 1) {
            // The chunk's last tile: its K-1 halo columns come from the
            // snapshot (the next chunk may already have overwritten them).
            const __hip_bfloat16 *hs = halo + (i64)blockIdx.x * C * (K - 1);
            for (int e = tid; e < C * (K - 1); e += NT) {
                const int ic = e / (K - 1), row = TO + e % (K - 1);
                *reinterpret_cast<__hip_bfloat16 *>(cur + row * RSB + 16 * ((ic >> 3) ^ ((row >> 2) & 15)) +
                                                    2 * (ic & 7)) = hs[e];
            }
            __syncthreads();
        }
        // The last tile of a chunk stages itself again into the unused
        // buffer, so that the block code has no branch around the staging.
        const int stage_tl = tile + 1 < tile_end ? tile + 1 : tile;
        const __amdgpu_buffer_rsrc_t rs_ld = rsrc_of(stage_tl);
        const int tcol_ld = (stage_tl % tiles) * TO;
        const __amdgpu_buffer_rsrc_t rs_st = rsrc_of(tile);
        const int tcol = t * TO;
        const int lim = Lout - tcol;  // block c0 stores local columns c0 + n < min(OB, lim - c0)

        // B-fragment addressing of block blk: row c0 + n, chunk (2s + h) ^ M(row).
        // The lane's column is re-read per tile (an opaque copy), so that the
        // compiler computes these addresses per block instead of holding all
        // of them in registers across the loop.
        int n_t = n;
        asm volatile("" : "+v"(n_t));
        auto b_addr = [&](int blk, int &base, int &sw) __attribute__((always_inline)) {
            const int row = blk * OB + n_t;
            const int m = (row >> 2) & 15;
            base = (it & 1) * BUFB + row * RSB + 16 * (h ^ (m & 1));
            sw = 32 * (m >> 1);
        };
        auto b_read = [&](int base, int sw, int s) __attribute__((always_inline)) {
            return *reinterpret_cast<const bf16x8_t *>(lds_raw + base + ((32 * s) ^ sw));
        };
        int b_base, b_sw, nb_base = 0, nb_sw = 0;
        b_addr(0, b_base, b_sw);
        bf16x8_t fr[3];
        fr[0] = b_read(b_base, b_sw, 0);
        fr[1] = b_read(b_base, b_sw, 1);
        u32x2_t ld[2][8];

        static_for<0, TB>([&](auto blk_) __attribute__((always_inline)) {
            constexpr int BLK = decltype(blk_)::value;
            f32x16_t(&cur_acc)[K] = (BLK & 1) ? acc1 : acc0;
            f32x16_t(&prv)[K] = (BLK & 1) ? acc0 : acc1;
            // the previous block: this tile's block BLK-1, or the previous tile's last
            constexpr int c0p = BLK ? (BLK - 1) * OB : (TB - 1) * OB;
            const __amdgpu_buffer_rsrc_t rsp = BLK ? rs_st : rs_prev;
            const int tcp = BLK ? tcol : tcol_prev;
            const int limp = BLK ? lim : lim_prev;
            u32x4_t sec_v[2];
            static_for<0, KS>([&](auto s_) __attribute__((always_inline)) {
                constexpr int S = decltype(s_)::value;
                constexpr int F = BLK * KS + S;  // fragment index within the tile
                // read two k-steps ahead (the next block's first two in this block's last two)
                if constexpr (F + 2 < TB * KS) {
                    if constexpr (S + 2 < KS)
                        fr[(F + 2) % 3] = b_read(b_base, b_sw, S + 2);
                    else
                        fr[(F + 2) % 3] = b_read(nb_base, nb_sw, S + 2 - KS);
                }
                const bf16x8_t bx = fr[F % 3];
                static_for<0, K>([&](auto k_) __attribute__((always_inline)) {
                    constexpr int KK = decltype(k_)::value;
                    if constexpr (S == 0)
                        asm volatile("v_mfma_f32_32x32x16_bf16 %0, %1, %2, 0"
                                     : "=&v"(cur_acc[KK])
                                     : "a"(wf[KK * KS]), "v"(bx));
                    else
                        asm volatile("v_mfma_f32_32x32x16_bf16 %0, %1, %2, %0"
                                     : "+v"(cur_acc[KK])
                                     : "a"(wf[KK * KS + S]), "v"(bx));
                    // ---- one slice of the other work ----
                    if constexpr (S == 0 && KK == 0) {
                        // fence: the previous block's accumulators are read by VALU below
                        asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 3" : "+v"(prv[0]), "+v"(prv[1]));
#pragma unroll
                        for (int q = 2; q < K; ++q) asm volatile("" : "+v"(prv[q]));
                    } else if constexpr (S == 0 && KK == 1) {
                        if constexpr (BLK + 1 < TB) b_addr(BLK + 1, nb_base, nb_sw);
                    } else if constexpr (S >= 1 && S <= K - 1) {
                        // Horner level: prv[kk] += S(prv[kk + 1]), four registers per MFMA
                        constexpr int KKL = K - 1 - S;
                        if constexpr (KK < 4) horner(prv[KKL], prv[KKL + 1], 4 * KK, 4);
                    } else if constexpr (S == K) {
                        constexpr int PB = BLK ? BLK - 1 : TB - 1;  // the previous block's index in its tile
                        if constexpr (KK < 4) {
                            ring_pair(prv[0], 4 * KK, c0p);
                            ring_pair(prv[0], 4 * KK + 2, c0p);
                        } else if constexpr (PB >= 1) {
                            sec_v[0] = ring_read(PB - 1, 0);
                            sec_v[1] = ring_read(PB - 1, 1);
                        }
                    } else if constexpr (S == K + 1) {
                        // staging store of the task loaded two blocks ago
                        static_for<0, TPT>([&](auto u_) __attribute__((always_inline)) {
                            constexpr int U = decltype(u_)::value;
                            if constexpr (tap_ld_blk<TB>(U) + 2 == BLK && KK < 4) stage_store_col(nxt, U, KK, ld[U & 1]);
                        });
                    } else if constexpr (S == K + 2) {
                        static_for<0, TPT>([&](auto u_) __attribute__((always_inline)) {
                            constexpr int U = decltype(u_)::value;
                            if constexpr (tap_ld_blk<TB>(U) == BLK && KK == 0) stage_load(rs_ld, tcol_ld, U, ld[U & 1]);
                        });
                        constexpr int PB = BLK ? BLK - 1 : TB - 1;
                        if constexpr (PB >= 1 && KK >= 1 && KK <= 2) store_sector(sec_v[KK - 1], PB - 1, KK - 1, rsp, tcp, limp);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            });
            if constexpr (BLK + 1 < TB) {
                b_base = nb_base;
                b_sw = nb_sw;
            }
        });
        __syncthreads();
        rs_prev = rs_st;
        tcol_prev = tcol;
        lim_prev = lim;
        tile = tile + 1;
    }
    // The last block's combination and stores (accumulator set 1: TB is even).
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 3" : "+v"(acc1[0]), "+v"(acc1[1]));
#pragma unroll
    for (int k = 2; k < K; ++k) asm volatile("" : "+v"(acc1[k]));
#pragma unroll
    for (int kk = K - 2; kk >= 0; --kk) horner(acc1[kk], acc1[kk + 1], 0, 16);
#pragma unroll
    for (int r = 0; r < 16; r += 2) ring_pair(acc1[0], r, (TB - 1) * OB);
#pragma unroll
    for (int i = 0; i < 2; ++i) store_sector(ring_read(TB - 2, i), TB - 2, i, rs_prev, tcol_prev, lim_prev);
}

// Three-term split W fragments: w = w0 + w1 + w2 exactly (each bf16 piece
// takes the next 8 significant bits of the FP32 weight), in the layout of
// conv_pack_weights_kernel, planes [3][C/32][KS][64][8].
__global__ void conv_pack_weights3_kernel(const float *__restrict__ w, int C, int K, __bf16 *__restrict__ out) {
    const int KS = C * K / 16;
    const i64 plane = (i64)(C / 32) * KS * 64 * 8;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < plane; idx += (i64)gridDim.x * blockDim.x) {
        const int e = (int)(idx % 8);
        const int l = (int)((idx / 8) % 64);
        const int s = (int)((idx / 512) % KS);
        const int wv = (int)(idx / (512 * (i64)KS));
        const int oc = 32 * wv + l % 32;
        const int r = 16 * s + 8 * (l / 32) + e;
        const int k = r / C, ic = r - k * C;
        const float v = w[((i64)oc * C + ic) * K + k];
        const __bf16 p0 = to_bf16(v);
        const float r1 = v - (float)p0;  // exact
        const __bf16 p1 = to_bf16(r1);
        out[idx] = p0;
        out[plane + idx] = p1;
        out[2 * plane + idx] = to_bf16(r1 - (float)p1);
    }
}

// FP32-accurate conv from bf16 MFMAs: x and w are split into three bf16
// terms (x0 + x1 + x2 and w0 + w1 + w2 represent the FP32 values exactly)
// and the six products down to 2^-16 relative are accumulated in FP32:
// w0 x0 into one accumulator, w0 x1 + w1 x0 + w0 x2 + w1 x1 + w2 x0 into a
// second, so their rounding errors stay 2^-8 below the leading term's.  The
// W fragments (3 x 16 B per lane per k-step) stream from L2 one k-step ahead;
// the three x planes are staged in LDS, double-buffered, as in
// conv_fp8_kernel.  FP32 storage only.
template <int C, int K, int TN, bool VL>
__global__ __launch_bounds__(2 * C, 1) void conv_bf16x6_kernel(float *__restrict__ x, const float *__restrict__ halo,
                                                             const bf16x8_t *__restrict__ wq,
                                                             const float *__restrict__ bias, int L, int tiles,
                                                             int total_tiles, int chunk) {
    constexpr int NW = C / 32;
    constexpr int NT = 64 * NW;
    constexpr int KS = C * K / 16;
    constexpr int JW = TN + K - 1;
    constexpr int RS = C + kConvPad;  // bf16 elements per LDS row
    constexpr int BUF = JW * RS;  // one plane of one buffer
    constexpr int NSUB = TN / 32;
    constexpr int NQ = TN / 4;
    constexpr int NTASK = (C / 8) * NQ;
    static_assert(NTASK % NT == 0, "staging tasks must divide evenly");
    constexpr int TPT = NTASK / NT;
    static_assert(KS % TPT == 0, "k-steps must split evenly over the staging tasks");
    constexpr int PH = KS / TPT;
    constexpr int NH = C * (K - 1);
    constexpr int HPT = (NH + NT - 1) / NT;
    constexpr i64 WPLANE = (i64)NW * KS * 64;  // fragments per weight plane
    using V4 = typename IoVec<float, 4>::type;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16 *xs = reinterpret_cast<__bf16 *>(lds_raw);  // [2][3][JW][RS]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int Lout = L - K + 1;

    int tile = blockIdx.x * chunk;
    const int tile_end = min(tile + chunk, total_tiles);
    if (tile >= tile_end) return;  // empty chunk (uniform over the work-group)

    const bf16x8_t *wsrc = wq + ((i64)wv * KS) * 64 + lane;
    V4 pf[8];
    float ph[HPT];
    auto load_task = [&](int tl, int u) __attribute__((always_inline)) {
        const int b = tl / tiles, t = tl - b * tiles;
        const int task = tid + u * NT;
        const int q = task % NQ, g = task / NQ;
        const float *xb = x + (i64)b * C * L;
#pragma unroll
        for (int i = 0; i < 8; ++i) load_quad_vec<float, VL>(xb + (i64)(8 * g + i) * L, t * TN + 4 * q, L, pf[i]);
    };
    auto store_task = [&](int buf, int tl, int u) __attribute__((always_inline)) {
        __bf16 *p0 = xs + (buf * 3) * BUF;
        const int task = tid + u * NT;
        const int q = task % NQ, g = task / NQ;
        const int j = (tl % tiles) * TN + 4 * q;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool in = j + c < L;
            bf16x8_t h0, h1, h2;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float v = in ? pf[i][c] : 0.0f;
                const __bf16 a = to_bf16(v);
                const float r1 = v - (float)a;
                const __bf16 bq = to_bf16(r1);
                h0[i] = a;
                h1[i] = bq;
                h2[i] = to_bf16(r1 - (float)bq);
            }
            const int off = (4 * q + c) * RS + 8 * g;
            *reinterpret_cast<bf16x8_t *>(p0 + off) = h0;
            *reinterpret_cast<bf16x8_t *>(p0 + BUF + off) = h1;
            *reinterpret_cast<bf16x8_t *>(p0 + 2 * BUF + off) = h2;
        }
    };
    auto load_halo = [&](int tl) __attribute__((always_inline)) {
        const int b = tl / tiles, t = tl - b * tiles;
        const float *xb = x + (i64)b * C * L;
        const bool snap = tl == tile_end - 1 && t < tiles - 1;
#pragma unroll
        for (int u = 0; u < HPT; ++u) {
            const int e = min(tid + u * NT, NH - 1);
            const int ic = e / (K - 1), kk = e - ic * (K - 1);
            const int j = t * TN + TN + kk;
            ph[u] = *(snap ? halo + (i64)blockIdx.x * NH + e : xb + (i64)ic * L + (j < L ? j : L - 1));
        }
    };
    auto store_halo = [&](int buf, int tl) __attribute__((always_inline)) {
        __bf16 *p0 = xs + (buf * 3) * BUF;
        const int t = tl % tiles;
        const bool snap = tl == tile_end - 1 && t < tiles - 1;
#pragma unroll
        for (int u = 0; u < HPT; ++u) {
            const int e = tid + u * NT;
            const int ic = e / (K - 1), kk = e - ic * (K - 1);
            const float v = snap || t * TN + TN + kk < L ? ph[u] : 0.0f;
            const __bf16 a = to_bf16(v);
            const float r1 = v - (float)a;
            const __bf16 bq = to_bf16(r1);
            const int off = (TN + kk) * RS + ic;
            if (e < NH) {
                p0[off] = a;
                p0[BUF + off] = bq;
                p0[2 * BUF + off] = to_bf16(r1 - (float)bq);
            }
        }
    };

#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        load_task(tile, u);
        store_task(0, tile, u);
    }
    load_halo(tile);
    store_halo(0, tile);
    __syncthreads();
    const int jl = lane % 32, kh = 8 * (lane / 32);
    for (int it = 0; tile < tile_end; ++it) {
        const int next = tile + 1;
        const bool has_next = next < tile_end;
        const __bf16 *p0 = xs + ((it & 1) * 3) * BUF;
        f32x16_t acc[NSUB], acl[NSUB];
#pragma unroll
        for (int n = 0; n < NSUB; ++n)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[n][r] = acl[n][r] = 0.0f;
        if (has_next) load_halo(next);
        bf16x8_t a0 = wsrc[0], a1 = wsrc[WPLANE], a2 = wsrc[2 * WPLANE];
        for (int u = 0; u < TPT; ++u) {
            if (has_next) load_task(next, u);
#pragma unroll 1
            for (int ss = 0; ss < PH; ++ss) {
                const int s = u * PH + ss;
                const i64 sn = ((s + 1) % KS) * 64;  // next k-step (wraps to the next tile's first)
                const bf16x8_t n0 = wsrc[sn], n1 = wsrc[WPLANE + sn], n2 = wsrc[2 * WPLANE + sn];
                const int r0 = 16 * s;
                const int k = r0 / C, ic0 = r0 - k * C + kh;
#pragma unroll
                for (int n = 0; n < NSUB; ++n) {
                    const int off = (32 * n + jl + k) * RS + ic0;
                    const bf16x8_t b0 = *reinterpret_cast<const bf16x8_t *>(p0 + off);
                    const bf16x8_t b1 = *reinterpret_cast<const bf16x8_t *>(p0 + BUF + off);
                    const bf16x8_t b2 = *reinterpret_cast<const bf16x8_t *>(p0 + 2 * BUF + off);
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[n], 0, 0, 0);
                    acl[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acl[n], 0, 0, 0);
                    acl[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acl[n], 0, 0, 0);
                    acl[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b2, acl[n], 0, 0, 0);
                    acl[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acl[n], 0, 0, 0);
                    acl[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b0, acl[n], 0, 0, 0);
                }
                a0 = n0;
                a1 = n1;
                a2 = n2;
            }
            if (has_next) store_task((it + 1) & 1, next, u);
        }
        if (has_next) store_halo((it + 1) & 1, next);
        {
            const int b = tile / tiles, t = tile - b * tiles;
            float *xb = x + (i64)b * C * L;
#pragma unroll
            for (int n = 0; n < NSUB; ++n) {
                const int j = t * TN + 32 * n + jl;
                if (j < Lout) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int oc = 32 * wv + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
                        xb[(i64)oc * L + j] = (acc[n][r] + acl[n][r]) + bias[oc];
                    }
                }
            }
        }
        __syncthreads();
        tile = next;
    }
}

// ---------------------------------------------------------------------------
// FP8 (OCP e4m3) kernels
// ---------------------------------------------------------------------------

// Per-output-channel weight scale: the power of two 2^e >= amax/448 (so that
// |w/scale| <= 448 and integers with few significant bits stay exact).
__global__ void conv_fp8_weight_scale_kernel(const float *__restrict__ w, int C, int K, float *__restrict__ scale) {
    const int oc = blockIdx.x;
    __shared__ float red[256];
    float m = 0.0f;
    for (int i = threadIdx.x; i < C * K; i += blockDim.x) m = fmaxf(m, fabsf(w[(i64)oc * C * K + i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float amax = red[0];
        scale[oc] = amax > 0.0f ? exp2f(ceilf(log2f(amax / 448.0f))) : 1.0f;
    }
}

// FP8 W fragments: wave w (channels 32w..), k-step s (r in [64s, 64s+64)),
// lane l: 32 bytes A[oc = 32w + l%32][r = 64s + 32*(l/32) + e], r = k*C + ic,
// values w / scale[oc].
__global__ void conv_fp8_pack_weights_kernel(const float *__restrict__ w, const float *__restrict__ scale, int C,
                                             int K, unsigned char *__restrict__ out) {
    const int KS = C * K / 64;
    const i64 total = (i64)(C / 32) * KS * 64 * 16;  // byte pairs
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const int e2 = (int)(idx % 16);
        const int l = (int)((idx / 16) % 64);
        const int s = (int)((idx / 1024) % KS);
        const int wv = (int)(idx / (1024 * (i64)KS));
        const int oc = 32 * wv + l % 32;
        float v[2];
        for (int h = 0; h < 2; ++h) {
            const int r = 64 * s + 32 * (l / 32) + 2 * e2 + h;
            const int k = r / C, ic = r - k * C;
            v[h] = w[((i64)oc * C + ic) * K + k] / scale[oc];
        }
        const int p = pack_fp8x2(v[0], v[1]);
        out[2 * idx] = (unsigned char)(p & 0xff);
        out[2 * idx + 1] = (unsigned char)((p >> 8) & 0xff);
    }
}

// Persistent FP8 implicit GEMM.  A work-group (C/32 waves) owns all C output
// channels of a tile of TN columns; wave w computes channels 32w..32w+31 and
// streams its W fragments (FP8, 32 bytes per lane per k-step of 64) from
// global memory (L2-resident), one k-step ahead.  The input tile is staged
// in LDS as FP8 e4m3 (x / xscale), channel-last, double-buffered; the next
// tile's global loads, conversion and LDS stores are spread over the k-loop
// of the current one.  y = acc * wscale[oc] * xscale + bias[oc].
template <class TIO, int C, int K, int TN, bool VL>
__global__ __launch_bounds__(2 * C, 1) void conv_fp8_kernel(TIO *__restrict__ x, const TIO *__restrict__ halo,
                                                          const i32x8_t *__restrict__ wq,
                                                          const float *__restrict__ wscale,
                                                          const float *__restrict__ bias, float xscale, int L,
                                                          int tiles, int total_tiles, int chunk) {
    constexpr int NW = C / 32;
    constexpr int NT = 64 * NW;
    constexpr int KS = C * K / 64;
    constexpr int JW = TN + K - 1;
    constexpr int RS = C + 16;  // bytes per LDS row
    constexpr int BUF = JW * RS;
    constexpr int NSUB = TN / 32;
    constexpr int NQ = TN / 4;
    constexpr int NTASK = (C / 8) * NQ;
    static_assert(NTASK % NT == 0, "staging tasks must divide evenly");
    constexpr int TPT = NTASK / NT;
    static_assert(KS % TPT == 0, "k-steps must split evenly over the staging tasks");
    constexpr int PH = KS / TPT;  // k-steps per staging task
    constexpr int NH = C * (K - 1);
    static_assert(NH % NT == 0, "halo elements must divide evenly");
    constexpr int HPT = NH / NT;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int Lout = L - K + 1;
    const float inv_xs = 1.0f / xscale;

    int tile = blockIdx.x * chunk;
    const int tile_end = min(tile + chunk, total_tiles);
    if (tile >= tile_end) return;  // empty chunk (uniform over the work-group)

    const i32x8_t *wsrc = wq + ((i64)wv * KS) * 64 + lane;

    TIO pf[8][4];
    TIO ph[HPT];
    auto load_task = [&](int tl, int u) __attribute__((always_inline)) {
        const int b = tl / tiles, t = tl - b * tiles;
        const int task = tid + u * NT;
        const int q = task % NQ, g = task / NQ;
        const int j = t * TN + 4 * q;
        const TIO *xb = x + (i64)b * C * L;
#pragma unroll
        for (int i = 0; i < 8; ++i) load_quad_raw<TIO, VL>(xb + (i64)(8 * g + i) * L, j, L, pf[i]);
    };
    auto store_task = [&](int buf, int tl, int u) __attribute__((always_inline)) {
        unsigned char *xs = lds_raw + buf * BUF;
        const int task = tid + u * NT;
        const int q = task % NQ, g = task / NQ;
        const int j = (tl % tiles) * TN + 4 * q;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float sc = j + c < L ? inv_xs : 0.0f;  // zero past L
            i32x2_t v;
            v[0] = (pack_fp8x2(io_to_f32(pf[0][c]) * sc, io_to_f32(pf[1][c]) * sc) & 0xffff) |
                   (pack_fp8x2(io_to_f32(pf[2][c]) * sc, io_to_f32(pf[3][c]) * sc) << 16);
            v[1] = (pack_fp8x2(io_to_f32(pf[4][c]) * sc, io_to_f32(pf[5][c]) * sc) & 0xffff) |
                   (pack_fp8x2(io_to_f32(pf[6][c]) * sc, io_to_f32(pf[7][c]) * sc) << 16);
            *reinterpret_cast<i32x2_t *>(xs + (4 * q + c) * RS + 8 * g) = v;
        }
    };
    auto load_halo = [&](int tl) __attribute__((always_inline)) {
        const int b = tl / tiles, t = tl - b * tiles;
        const TIO *xb = x + (i64)b * C * L;
        const bool snap = tl == tile_end - 1 && t < tiles - 1;
#pragma unroll
        for (int u = 0; u < HPT; ++u) {
            const int e = tid + u * NT;
            const int ic = e / (K - 1), kk = e - ic * (K - 1);
            const int j = t * TN + TN + kk;
            ph[u] = *(snap ? halo + (i64)blockIdx.x * NH + e : xb + (i64)ic * L + (j < L ? j : L - 1));
        }
    };
    auto store_halo = [&](int buf, int tl) __attribute__((always_inline)) {
        unsigned char *xs = lds_raw + buf * BUF;
        const int t = tl % tiles;
        const bool snap = tl == tile_end - 1 && t < tiles - 1;
#pragma unroll
        for (int u = 0; u < HPT; ++u) {
            const int e = tid + u * NT;
            const int ic = e / (K - 1), kk = e - ic * (K - 1);
            const float sc = snap || t * TN + TN + kk < L ? inv_xs : 0.0f;
            xs[(TN + kk) * RS + ic] = (unsigned char)(pack_fp8x2(io_to_f32(ph[u]) * sc, 0.0f) & 0xff);
        }
    };

#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        load_task(tile, u);
        store_task(0, tile, u);
    }
    load_halo(tile);
    store_halo(0, tile);
    __syncthreads();
    const int jl = lane % 32, kh = 32 * (lane / 32);
    for (int it = 0; tile < tile_end; ++it) {
        const int next = tile + 1;
        const bool has_next = next < tile_end;
        const unsigned char *xs = lds_raw + (it & 1) * BUF;
        f32x16_t acc[NSUB];
#pragma unroll
        for (int n = 0; n < NSUB; ++n)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[n][r] = 0.0f;
        if (has_next) load_halo(next);
        i32x8_t wcur = wsrc[0];
        for (int u = 0; u < TPT; ++u) {
            if (has_next) load_task(next, u);
#pragma unroll 1
            for (int ss = 0; ss < PH; ++ss) {
                const int s = u * PH + ss;
                const i32x8_t wnext = wsrc[((s + 1) % KS) * 64];  // next k-step (wraps to the next tile's first)
                const int r0 = 64 * s;
                const int k = r0 / C, ic0 = r0 - k * C + kh;
                const unsigned char *src0 = xs + (jl + k) * RS + ic0;
#pragma unroll
                for (int n = 0; n < NSUB; ++n) {
                    const unsigned char *src = src0 + 32 * n * RS;
                    const i32x4_t b0 = *reinterpret_cast<const i32x4_t *>(src);
                    const i32x4_t b1 = *reinterpret_cast<const i32x4_t *>(src + 16);
                    const i32x8_t bv = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
                    acc[n] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(wcur, bv, acc[n], 0, 0, 0, 127, 0, 127);
                }
                wcur = wnext;
            }
            if (has_next) store_task((it + 1) & 1, next, u);
        }
        if (has_next) store_halo((it + 1) & 1, next);
        {
            const int b = tile / tiles, t = tile - b * tiles;
            TIO *xb = x + (i64)b * C * L;
#pragma unroll
            for (int n = 0; n < NSUB; ++n) {
                const int j = t * TN + 32 * n + jl;
                if (j < Lout) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        // scales and bias re-read here (L1-resident) rather than held in registers
                        const int oc = 32 * wv + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
                        xb[(i64)oc * L + j] = f32_to_io<TIO>(acc[n][r] * (wscale[oc] * xscale) + bias[oc]);
                    }
                }
            }
        }
        __syncthreads();
        tile = next;
    }
}


// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct ConvPlan {
    bkma_conv_mode mode;
    bkma_dtype io;
    i64 B, C, L, K;
    void *w = nullptr;      // direct modes: weights in io precision; MFMA: bf16 hi fragments
    void *wlo = nullptr;    // BF16X3: bf16 lo fragments
    void *bias = nullptr;   // direct: io precision; MFMA: f32
    void *halo = nullptr;   // MFMA: tile-boundary columns, snapshotted before each run
    void *wscale = nullptr; // FP8: per-output-channel weight scales (f32)
    float xscale = 1.0f;    // FP8: the input is quantized as x / xscale
    i64 resident_groups = 0;
    u64 lds_bytes = 0;
    std::string name;
    ~ConvPlan() {
        if (w) (void)hipFree(w);
        if (wlo) (void)hipFree(wlo);
        if (bias) (void)hipFree(bias);
        if (halo) (void)hipFree(halo);
        if (wscale) (void)hipFree(wscale);
    }
};

template <class Dst, class Src> __global__ void convert_kernel(Dst *dst, const Src *src, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = (Dst)src[i];
}

using MfmaLaunch = void (*)(void *x, const ConvPlan &p, hipStream_t s);

template <class TIO>
void launch_halo(void *x, const ConvPlan &p, int C, int K, int TN, int tiles, int total, int *grid, int *chunk,
                 hipStream_t s);

template <class TIO, int NPROD, int C, int K, int TN> void launch_mfma(void *x, const ConvPlan &p, hipStream_t s) {
    const int tiles = (int)ceil_div((u64)(p.L - p.K + 1), (u64)TN);
    const int total = (int)(p.B * tiles);
    int grid = 0, chunk = 0;
    launch_halo<TIO>(x, p, C, K, TN, tiles, total, &grid, &chunk, s);
    auto k = p.L % 4 == 0 ? conv_mfma_kernel<TIO, NPROD, C, K, TN, true> : conv_mfma_kernel<TIO, NPROD, C, K, TN, false>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_bytes);
    hipLaunchKernelGGL(k, dim3(grid), dim3(2 * C), (size_t)p.lds_bytes, s, reinterpret_cast<TIO *>(x),
                       reinterpret_cast<const TIO *>(p.halo), (const __bf16 *)p.w, (const __bf16 *)p.wlo,
                       (const float *)p.bias, (int)p.L, tiles, total, chunk);
}

template <class TIO, int NPROD, int C, int K, int TN> int occupancy_mfma(u64 lds_bytes) {
    auto k = conv_mfma_kernel<TIO, NPROD, C, K, TN, true>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 2 * C, (size_t)lds_bytes) != hipSuccess) return 0;
    return blocks;
}

// Tap-reuse kernel (BF16 storage): TB blocks of 33-K output columns per tile.
template <int C, int K, int TB> void launch_tap(void *x, const ConvPlan &p, hipStream_t s) {
    constexpr int TO = TB * (33 - K);
    const int tiles = (int)ceil_div((u64)(p.L - p.K + 1), (u64)TO);
    const int total = (int)(p.B * tiles);
    int grid = 0, chunk = 0;
    launch_halo<__hip_bfloat16>(x, p, C, K, TO, tiles, total, &grid, &chunk, s);
    auto k = conv_tap_bf16_kernel<C, K, TB>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_bytes);
    hipLaunchKernelGGL(k, dim3(grid), dim3(2 * C), (size_t)p.lds_bytes, s, reinterpret_cast<__hip_bfloat16 *>(x),
                       reinterpret_cast<const __hip_bfloat16 *>(p.halo), (const __bf16 *)p.w, (const float *)p.bias,
                       (int)p.L, tiles, total, chunk);
}

template <int C, int K, int TB> int occupancy_tap(u64 lds_bytes) {
    auto k = conv_tap_bf16_kernel<C, K, TB>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 2 * C, (size_t)lds_bytes) != hipSuccess) return 0;
    return blocks;
}

// TN: output columns per tile; tap: the tap-reuse kernel (TN = TB * (33 - K)).
struct MfmaVariant {
    bkma_dtype io;
    int nprod, C, K, TN;
    MfmaLaunch fn;
    int (*occ)(u64 lds_bytes);
    int tap = 0;
};

#define MV(TIO, IO, NP, C, K, TN) {IO, NP, C, K, TN, &launch_mfma<TIO, NP, C, K, TN>, &occupancy_mfma<TIO, NP, C, K, TN>, 0}
#define TV(C, K, TB) {BKMA_BF16, 1, C, K, TB * (33 - K), &launch_tap<C, K, TB>, &occupancy_tap<C, K, TB>, 1}
// The first entry that matches is taken; the tap kernel needs L % 8 == 0 and
// C*L*2 < 2^31, the entries after it cover the other lengths.
const MfmaVariant kMfmaVariants[] = {
    TV(128, 5, 8),
    MV(__hip_bfloat16, BKMA_BF16, 1, 128, 5, 128), MV(__hip_bfloat16, BKMA_BF16, 3, 128, 5, 64),
    MV(float, BKMA_F32, 1, 128, 5, 128),           MV(float, BKMA_F32, 3, 128, 5, 64),
    MV(float, BKMA_F32, 1, 128, 3, 128),      MV(float, BKMA_F32, 3, 128, 3, 64),
    MV(float, BKMA_F32, 1, 64, 5, 128),       MV(float, BKMA_F32, 3, 64, 5, 64),
};
#undef MV
#undef TV

// Grid (one work-group per chunk of consecutive tiles) and the halo copies
// of the chunks' last tiles.
template <class TIO>
void launch_halo(void *x, const ConvPlan &p, int C, int K, int TN, int tiles, int total, int *grid, int *chunk,
                 hipStream_t s) {
    const int g = (int)std::max<i64>(1, std::min<i64>(total, p.resident_groups));
    *chunk = (int)ceil_div((u64)total, (u64)g);
    *grid = (int)ceil_div((u64)total, (u64)*chunk);
    if (tiles > 1)
        hipLaunchKernelGGL(conv_halo_kernel<TIO>, dim3(*grid), dim3(256), 0, s, reinterpret_cast<const TIO *>(x),
                           reinterpret_cast<TIO *>(p.halo), C, (int)p.L, K, TN, tiles, total, *chunk);
}

template <class TIO, int C, int K, int TN> void launch_fp8(void *x, const ConvPlan &p, hipStream_t s) {
    const int tiles = (int)ceil_div((u64)(p.L - p.K + 1), (u64)TN);
    const int total = (int)(p.B * tiles);
    int grid = 0, chunk = 0;
    launch_halo<TIO>(x, p, C, K, TN, tiles, total, &grid, &chunk, s);
    auto k = p.L % 4 == 0 ? conv_fp8_kernel<TIO, C, K, TN, true> : conv_fp8_kernel<TIO, C, K, TN, false>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_bytes);
    hipLaunchKernelGGL(k, dim3(grid), dim3(2 * C), (size_t)p.lds_bytes, s, reinterpret_cast<TIO *>(x),
                       reinterpret_cast<const TIO *>(p.halo), (const i32x8_t *)p.w, (const float *)p.wscale,
                       (const float *)p.bias, p.xscale, (int)p.L, tiles, total, chunk);
}

template <class TIO, int C, int K, int TN> int occupancy_fp8(u64 lds_bytes) {
    auto k = conv_fp8_kernel<TIO, C, K, TN, true>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 2 * C, (size_t)lds_bytes) != hipSuccess) return 0;
    return blocks;
}

#define FV(TIO, IO, C, K, TN) {IO, 0, C, K, TN, &launch_fp8<TIO, C, K, TN>, &occupancy_fp8<TIO, C, K, TN>}
const MfmaVariant kFp8Variants[] = {
    FV(float, BKMA_F32, 256, 7, 256),
    FV(__hip_bfloat16, BKMA_BF16, 256, 7, 256),
    FV(float, BKMA_F32, 128, 5, 128),
};
#undef FV

template <int C, int K, int TN> void launch_bf16x6(void *x, const ConvPlan &p, hipStream_t s) {
    const int tiles = (int)ceil_div((u64)(p.L - p.K + 1), (u64)TN);
    const int total = (int)(p.B * tiles);
    int grid = 0, chunk = 0;
    launch_halo<float>(x, p, C, K, TN, tiles, total, &grid, &chunk, s);
    auto k = p.L % 4 == 0 ? conv_bf16x6_kernel<C, K, TN, true> : conv_bf16x6_kernel<C, K, TN, false>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_bytes);
    hipLaunchKernelGGL(k, dim3(grid), dim3(2 * C), (size_t)p.lds_bytes, s, reinterpret_cast<float *>(x),
                       reinterpret_cast<const float *>(p.halo), (const bf16x8_t *)p.w, (const float *)p.bias,
                       (int)p.L, tiles, total, chunk);
}

template <int C, int K, int TN> int occupancy_bf16x6(u64 lds_bytes) {
    auto k = conv_bf16x6_kernel<C, K, TN, true>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 2 * C, (size_t)lds_bytes) != hipSuccess) return 0;
    return blocks;
}

#define XV(C, K, TN) {BKMA_F32, 6, C, K, TN, &launch_bf16x6<C, K, TN>, &occupancy_bf16x6<C, K, TN>}
const MfmaVariant kBf16x6Variants[] = {
    XV(128, 5, 64),
    XV(128, 3, 64),
    XV(64, 5, 64),
};
#undef XV

static const MfmaVariant *find_bf16x6(bkma_dtype io, i64 C, i64 K) {
    for (const auto &v : kBf16x6Variants)
        if (v.io == io && v.C == C && v.K == K) return &v;
    return nullptr;
}

static const MfmaVariant *find_fp8(bkma_dtype io, i64 C, i64 K) {
    for (const auto &v : kFp8Variants)
        if (v.io == io && v.C == C && v.K == K) return &v;
    return nullptr;
}

static const MfmaVariant *find_mfma(bkma_dtype io, int nprod, i64 C, i64 K, i64 L) {
    for (const auto &v : kMfmaVariants)
        if (v.io == io && v.nprod == nprod && v.C == C && v.K == K &&
            (!v.tap || (L % 8 == 0 && C * L * 2 < (i64(1) << 31))))
            return &v;
    return nullptr;
}

}  // namespace bkma

using namespace bkma;

struct bkma_conv1d_plan : ConvPlan {};

extern "C" bkma_status bkma_conv1d_plan_create(bkma_conv_mode mode, bkma_dtype io_dtype, uint64_t B, uint64_t C,
                                               uint64_t L, uint64_t K, bkma_dtype w_dtype, const void *weight,
                                               const void *bias, void *stream, bkma_conv1d_plan **out) {
    if (!out || !weight || !bias) return fail(BKMA_ERR_INVALID_ARGUMENT, "conv1d plan: null argument");
    *out = nullptr;
    if (B == 0 || C == 0 || K == 0 || L < K) return fail(BKMA_ERR_INVALID_ARGUMENT, "conv1d plan: need L >= K >= 1");
    if (B > 0x7fffffffULL || C * L > 0x7fffffffULL)
        return fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: sizes beyond 32-bit indexing");
    BKMA_TRY(require_device());
    hipStream_t s = as_stream(stream);
    auto *p = new (std::nothrow) bkma_conv1d_plan();
    if (!p) return fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan allocation");
    p->mode = mode;
    p->io = io_dtype;
    p->B = (i64)B;
    p->C = (i64)C;
    p->L = (i64)L;
    p->K = (i64)K;
    auto bail = [&](bkma_status st) {
        delete p;
        return st;
    };
    const i64 nw = (i64)(C * C * K);
    u64 lds_max = 0;
    BKMA_TRY(lds_limit_bytes(&lds_max));
    if (mode == BKMA_CONV_F64 || mode == BKMA_CONV_F32) {
        const bkma_dtype want = mode == BKMA_CONV_F64 ? BKMA_F64 : BKMA_F32;
        if (io_dtype != want || w_dtype != want)
            return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: F64/F32 modes take data and weights of that type"));
        const size_t esz = want == BKMA_F64 ? 8 : 4;
        if (hipMalloc(&p->w, nw * esz) != hipSuccess || hipMalloc(&p->bias, C * esz) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: weight copy"));
        if (hipMemcpyAsync(p->w, weight, nw * esz, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(p->bias, bias, C * esz, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: weight copy"));
        p->lds_bytes = (u64)C * (u64)(kDirTN + K - 1) * esz;
        if (p->lds_bytes > lds_max) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: C too large for the LDS tile"));
        p->name = want == BKMA_F64 ? "conv_direct<f64>" : "conv_direct<f32>";
    } else if (mode == BKMA_CONV_BF16 || mode == BKMA_CONV_BF16X3) {
        const int nprod = mode == BKMA_CONV_BF16 ? 1 : 3;
        const MfmaVariant *mv = find_mfma(io_dtype, nprod, (i64)C, (i64)K, (i64)L);
        if (!mv)
            return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: no MFMA kernel compiled for C=" + std::to_string(C) +
                                                       ", K=" + std::to_string(K) + " and this io dtype"));
        if (w_dtype != BKMA_F32) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: MFMA modes take F32 weights"));
        const i64 frag = nw;  // C*C*K bf16 values
        if (hipMalloc(&p->w, frag * 2) != hipSuccess || (nprod == 3 && hipMalloc(&p->wlo, frag * 2) != hipSuccess) ||
            hipMalloc(&p->bias, C * 4) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: weight fragments"));
        hipLaunchKernelGGL(conv_pack_weights_kernel, dim3(256), dim3(256), 0, s, (const float *)weight, (int)C, (int)K,
                           (__bf16 *)p->w, (__bf16 *)p->wlo);
        if (hipMemcpyAsync(p->bias, bias, C * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: bias copy"));
        const bool xlo = nprod == 3 && io_dtype != BKMA_BF16;
        p->lds_bytes = mv->tap ? 2ull * (u64)(mv->TN + K - 1) * (u64)(2 * C + 32) + (u64)(C / 32) * 32 * 192
                               : 2ull * (u64)(mv->TN + K - 1) * (u64)(C + kConvPad) * 2 * (xlo ? 2 : 1);
        if (p->lds_bytes > lds_max) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: LDS tile too large"));
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: device query"));
        p->resident_groups = (i64)std::max(1, mv->occ(p->lds_bytes)) * cus;
        const u64 esz = io_dtype == BKMA_BF16 ? 2 : 4;
        if (hipMalloc(&p->halo, std::max<u64>((u64)p->resident_groups * C * (K - 1) * esz, 16)) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: halo buffer"));
        p->name = std::string(mv->tap ? "conv_tap_bf16" : "conv_mfma_bf16") + (nprod == 3 ? "x3" : "") + "<C" +
                  std::to_string(C) + ",K" + std::to_string(K) + ",TN" + std::to_string(mv->TN) +
                  (io_dtype == BKMA_BF16 ? ",io_bf16" : ",io_f32") + ">";
    } else if (mode == BKMA_CONV_BF16X6) {
        const MfmaVariant *mv = find_bf16x6(io_dtype, (i64)C, (i64)K);
        if (!mv)
            return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: no BF16X6 kernel compiled for C=" + std::to_string(C) +
                                                       ", K=" + std::to_string(K) + " and this io dtype (F32 only)"));
        if (w_dtype != BKMA_F32) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: BF16X6 takes F32 weights"));
        if (hipMalloc(&p->w, (size_t)nw * 2 * 3) != hipSuccess || hipMalloc(&p->bias, C * 4) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: weight fragments"));
        hipLaunchKernelGGL(conv_pack_weights3_kernel, dim3(256), dim3(256), 0, s, (const float *)weight, (int)C, (int)K,
                           (__bf16 *)p->w);
        if (hipMemcpyAsync(p->bias, bias, C * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: bias copy"));
        p->lds_bytes = 2ull * 3 * (u64)(mv->TN + K - 1) * (u64)(C + kConvPad) * 2;
        if (p->lds_bytes > lds_max) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: LDS tile too large"));
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: device query"));
        p->resident_groups = (i64)std::max(1, mv->occ(p->lds_bytes)) * cus;
        if (hipMalloc(&p->halo, std::max<u64>((u64)p->resident_groups * C * (K - 1) * 4, 16)) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: halo buffer"));
        p->name = std::string("conv_mfma_bf16x6<C") + std::to_string(C) + ",K" + std::to_string(K) + ",TN" +
                  std::to_string(mv->TN) + ",io_f32>";
    } else if (mode == BKMA_CONV_FP8) {
        const MfmaVariant *mv = find_fp8(io_dtype, (i64)C, (i64)K);
        if (!mv)
            return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: no FP8 kernel compiled for C=" + std::to_string(C) +
                                                       ", K=" + std::to_string(K) + " and this io dtype"));
        if (w_dtype != BKMA_F32) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: FP8 mode takes F32 weights"));
        if (hipMalloc(&p->w, (size_t)nw) != hipSuccess || hipMalloc(&p->wscale, C * 4) != hipSuccess ||
            hipMalloc(&p->bias, C * 4) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: weight fragments"));
        hipLaunchKernelGGL(conv_fp8_weight_scale_kernel, dim3((unsigned)C), dim3(256), 0, s, (const float *)weight,
                           (int)C, (int)K, (float *)p->wscale);
        hipLaunchKernelGGL(conv_fp8_pack_weights_kernel, dim3(256), dim3(256), 0, s, (const float *)weight,
                           (const float *)p->wscale, (int)C, (int)K, (unsigned char *)p->w);
        if (hipMemcpyAsync(p->bias, bias, C * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: bias copy"));
        p->lds_bytes = 2ull * (u64)(mv->TN + K - 1) * (u64)(C + 16);
        if (p->lds_bytes > lds_max) return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: LDS tile too large"));
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: device query"));
        p->resident_groups = (i64)std::max(1, mv->occ(p->lds_bytes)) * cus;
        const u64 esz = io_dtype == BKMA_BF16 ? 2 : 4;
        if (hipMalloc(&p->halo, std::max<u64>((u64)p->resident_groups * C * (K - 1) * esz, 16)) != hipSuccess)
            return bail(fail(BKMA_ERR_OUT_OF_MEMORY, "conv1d plan: halo buffer"));
        p->name = std::string("conv_fp8_e4m3<C") + std::to_string(C) + ",K" + std::to_string(K) + ",TN" +
                  std::to_string(mv->TN) + (io_dtype == BKMA_BF16 ? ",io_bf16" : ",io_f32") + ">";
    } else {
        return bail(fail(BKMA_ERR_UNSUPPORTED, "conv1d plan: unknown mode"));
    }
    if (hipGetLastError() != hipSuccess) return bail(fail(BKMA_ERR_DEVICE, "conv1d plan: preparation kernel"));
    *out = p;
    return BKMA_OK;
}

extern "C" bkma_status bkma_conv1d_plan_run(bkma_conv1d_plan *p, void *x, void *stream) {
    if (!p || !x) return fail(BKMA_ERR_INVALID_ARGUMENT, "conv1d run: null argument");
    hipStream_t s = as_stream(stream);
    switch (p->mode) {
        case BKMA_CONV_F64: {
            auto k = conv_direct_kernel<double>;
            BKMA_HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes));
            hipLaunchKernelGGL(k, dim3((unsigned)p->B), dim3(256), (size_t)p->lds_bytes, s, (double *)x,
                               (const double *)p->w, (const double *)p->bias, (int)p->C, (int)p->L, (int)p->K);
            break;
        }
        case BKMA_CONV_F32: {
            auto k = conv_direct_kernel<float>;
            BKMA_HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes));
            hipLaunchKernelGGL(k, dim3((unsigned)p->B), dim3(256), (size_t)p->lds_bytes, s, (float *)x,
                               (const float *)p->w, (const float *)p->bias, (int)p->C, (int)p->L, (int)p->K);
            break;
        }
        case BKMA_CONV_BF16:
        case BKMA_CONV_BF16X3: {
            const MfmaVariant *v = find_mfma(p->io, p->mode == BKMA_CONV_BF16 ? 1 : 3, p->C, p->K, p->L);
            if (!v) return fail(BKMA_ERR_UNSUPPORTED, "conv1d run: no MFMA kernel");
            v->fn(x, *p, s);
            break;
        }
        case BKMA_CONV_BF16X6: {
            const MfmaVariant *v = find_bf16x6(p->io, p->C, p->K);
            if (!v) return fail(BKMA_ERR_UNSUPPORTED, "conv1d run: no BF16X6 kernel");
            v->fn(x, *p, s);
            break;
        }
        case BKMA_CONV_FP8: {
            const MfmaVariant *v = find_fp8(p->io, p->C, p->K);
            if (!v) return fail(BKMA_ERR_UNSUPPORTED, "conv1d run: no FP8 kernel");
            v->fn(x, *p, s);
            break;
        }
        default:
            return fail(BKMA_ERR_UNSUPPORTED, "conv1d run: unknown mode");
    }
    return check_launch("conv1d kernel");
}

extern "C" bkma_status bkma_conv1d_plan_describe(const bkma_conv1d_plan *p, char *name, uint64_t name_len) {
    if (!p) return fail(BKMA_ERR_INVALID_ARGUMENT, "null plan");
    if (name && name_len) {
        std::strncpy(name, p->name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    return BKMA_OK;
}

extern "C" bkma_status bkma_conv1d_plan_set_input_scale(bkma_conv1d_plan *p, float scale) {
    if (!p) return fail(BKMA_ERR_INVALID_ARGUMENT, "null plan");
    if (!(scale > 0.0f) || !std::isfinite(scale)) return fail(BKMA_ERR_INVALID_ARGUMENT, "input scale must be positive");
    p->xscale = scale;
    return BKMA_OK;
}

extern "C" void bkma_conv1d_plan_destroy(bkma_conv1d_plan *p) { delete p; }
