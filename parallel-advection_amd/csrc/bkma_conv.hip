// Conv1d (textbook semantics, torch.nn.functional.conv1d with stride 1 and
// no padding) written in place over its input:
//   y[b][oc][j] = bias[oc] + sum_ic sum_k w[oc][ic][k] * x[b][ic][j + k],
//   j in [0, L-K+1) overwrites x[b][oc][j]; x[b][oc][L-K+1 .. L) is kept.
//
// In-place safety: y[b][:, J] overwrites x[b][:, J], which every output
// channel still needs, so one work-group owns one batch element b and walks
// its columns in increasing tiles [j0, j0+TN): a tile reads columns
// [j0, j0+TN+K-1) and writes [j0, j0+TN), so reads always run ahead of
// writes.
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
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"

namespace bkma {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }  // round to nearest even
template <class T> __device__ __forceinline__ float io_to_f32(T v) { return (float)v; }
template <> __device__ __forceinline__ float io_to_f32<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <class T> __device__ __forceinline__ T f32_to_io(float v) { return (T)v; }
template <> __device__ __forceinline__ __hip_bfloat16 f32_to_io<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

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

// LDS tile: Xs[jj][ic] bf16, row stride C + 8 (16-byte pad keeps the rows of
// a ds_read_b128 on different banks).
template <class TIO, int NPROD, int KS, int NW>
__global__ __launch_bounds__(64 * NW) void conv_mfma_kernel(TIO *__restrict__ x, const __bf16 *__restrict__ whi,
                                                           const __bf16 *__restrict__ wlo, const float *__restrict__ bias,
                                                           int L, int K) {
    constexpr int C = 32 * NW;     // channels (in = out)
    constexpr int TN = 64;         // columns per tile (two 32-column MFMA tiles)
    constexpr int RS = C + 8;      // LDS row stride (elements)
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16 *xs_hi = reinterpret_cast<__bf16 *>(lds_raw);
    const int JW = TN + K - 1;
    __bf16 *xs_lo = xs_hi + JW * RS;
    const int Lout = L - K + 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    TIO *xb = x + (i64)blockIdx.x * C * L;

    // Weight fragments of this wave, resident for the whole kernel.
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

    for (int j0 = 0; j0 < Lout; j0 += TN) {
        __syncthreads();  // the previous tile's LDS reads are done
        // Stage X[ic][j0 .. j0+JW) channel-last, rounded to bf16 (+ residual).
        for (int e = threadIdx.x; e < C * JW; e += 64 * NW) {
            const int ic = e / JW, jj = e - ic * JW;
            const int j = j0 + jj;
            const float v = j < L ? io_to_f32(xb[(i64)ic * L + j]) : 0.0f;
            const __bf16 h = to_bf16(v);
            xs_hi[jj * RS + ic] = h;
            if constexpr (NPROD == 3) xs_lo[jj * RS + ic] = to_bf16(v - (float)h);
        }
        __syncthreads();

        f32x16_t acc[2];
#pragma unroll
        for (int n = 0; n < 2; ++n)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[n][r] = 0.0f;
        const int jl = lane % 32, kh = 8 * (lane / 32);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int r0 = 16 * s;          // compile-time after unrolling
            const int k = r0 / C, ic0 = r0 - k * C + kh;
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                const int row = 32 * n + jl + k;
                const bf16x8_t bh = *reinterpret_cast<const bf16x8_t *>(xs_hi + row * RS + ic0);
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], bh, acc[n], 0, 0, 0);
                if constexpr (NPROD == 3) {
                    const bf16x8_t bl = *reinterpret_cast<const bf16x8_t *>(xs_lo + row * RS + ic0);
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], bl, acc[n], 0, 0, 0);
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wb[s], bh, acc[n], 0, 0, 0);
                }
            }
        }
        // D[m][n]: lane l, register r -> m = 8*(r/4) + 4*(l/32) + r%4, n = l%32.
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            const int j = j0 + 32 * n + jl;
            if (j < Lout) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int oc = 32 * wv + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
                    xb[(i64)oc * L + j] = f32_to_io<TIO>(acc[n][r] + bias_r[r]);
                }
            }
        }
    }
}

// Plain 32x32x16 bf16 MFMA over one fragment pair, for the layout test:
// a[l][e], b[l][e] are the lane fragments, d[l][r] the accumulator.
__global__ void mfma_layout_probe_kernel(const __bf16 *a, const __bf16 *b, float *d) {
    const int l = threadIdx.x;
    bf16x8_t av, bv;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        av[e] = a[l * 8 + e];
        bv[e] = b[l * 8 + e];
    }
    f32x16_t c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0.0f;
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) d[l * 16 + r] = c[r];
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
    u64 lds_bytes = 0;
    std::string name;
    ~ConvPlan() {
        if (w) (void)hipFree(w);
        if (wlo) (void)hipFree(wlo);
        if (bias) (void)hipFree(bias);
    }
};

template <class Dst, class Src> __global__ void convert_kernel(Dst *dst, const Src *src, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = (Dst)src[i];
}

using MfmaLaunch = void (*)(void *x, const ConvPlan &p, hipStream_t s);

template <class TIO, int NPROD, int KS, int NW> void launch_mfma(void *x, const ConvPlan &p, hipStream_t s) {
    auto k = conv_mfma_kernel<TIO, NPROD, KS, NW>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_bytes);
    hipLaunchKernelGGL(k, dim3((unsigned)p.B), dim3(64 * NW), (size_t)p.lds_bytes, s, reinterpret_cast<TIO *>(x),
                       (const __bf16 *)p.w, (const __bf16 *)p.wlo, (const float *)p.bias, (int)p.L, (int)p.K);
}

struct MfmaVariant {
    bkma_dtype io;
    int nprod, C, K;
    MfmaLaunch fn;
};

const MfmaVariant kMfmaVariants[] = {
    {BKMA_F32, 1, 128, 5, &launch_mfma<float, 1, 40, 4>},
    {BKMA_F32, 3, 128, 5, &launch_mfma<float, 3, 40, 4>},
    {BKMA_BF16, 1, 128, 5, &launch_mfma<__hip_bfloat16, 1, 40, 4>},
    {BKMA_BF16, 3, 128, 5, &launch_mfma<__hip_bfloat16, 3, 40, 4>},
    {BKMA_F32, 1, 128, 3, &launch_mfma<float, 1, 24, 4>},
    {BKMA_F32, 3, 128, 3, &launch_mfma<float, 3, 24, 4>},
    {BKMA_F32, 1, 64, 5, &launch_mfma<float, 1, 20, 2>},
    {BKMA_F32, 3, 64, 5, &launch_mfma<float, 3, 20, 2>},
};

static const MfmaVariant *find_mfma(bkma_dtype io, int nprod, i64 C, i64 K) {
    for (const auto &v : kMfmaVariants)
        if (v.io == io && v.nprod == nprod && v.C == C && v.K == K) return &v;
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
        if (!find_mfma(io_dtype, nprod, (i64)C, (i64)K))
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
        p->lds_bytes = (u64)(64 + K - 1) * (u64)(C + 8) * 2 * (nprod == 3 ? 2 : 1);
        p->name = std::string("conv_mfma_bf16") + (nprod == 3 ? "x3" : "") + "<C" + std::to_string(C) + ",K" +
                  std::to_string(K) + (io_dtype == BKMA_BF16 ? ",io_bf16" : ",io_f32") + ">";
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
            const MfmaVariant *v = find_mfma(p->io, p->mode == BKMA_CONV_BF16 ? 1 : 3, p->C, p->K);
            if (!v) return fail(BKMA_ERR_UNSUPPORTED, "conv1d run: no MFMA kernel");
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

extern "C" void bkma_conv1d_plan_destroy(bkma_conv1d_plan *p) { delete p; }

extern "C" bkma_status bkma_mfma_layout_probe(const void *a, const void *b, void *d, void *stream) {
    if (!a || !b || !d) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    BKMA_TRY(require_device());
    hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, as_stream(stream), (const __bf16 *)a,
                       (const __bf16 *)b, (float *)d);
    return check_launch("mfma_layout_probe_kernel");
}
