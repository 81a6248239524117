// Operand-layout and conversion probes of the MFMA Conv1d kernels: single
// MFMAs on lane fragments and the kernels' FP32 -> FP8 e4m3 conversion
// (parallel-advection_amd/csrc/mfma_ops.hpp).  TEST INFRASTRUCTURE: built into
// tests/cpp/libbkma_probes.so, called by tests/test_conv_gpu.py.
#include <algorithm>

#include "mfma_ops.hpp"

using namespace bkma;

namespace {

// One v_mfma_f32_32x32x16_bf16: a[l][e], b[l][e] are lane l's fragments, d[l][r] the accumulator.
__global__ void mfma_bf16_kernel(const __bf16 *a, const __bf16 *b, float *d) {
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

// One v_mfma_scale_f32_32x32x64_f8f6f4 (e4m3 operands, unit block scales E8M0 127 = 2^0).
__global__ void mfma_fp8_kernel(const i32x8_t *a, const i32x8_t *b, float *d) {
    const int l = threadIdx.x;
    f32x16_t c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0.0f;
    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[l], b[l], c, 0, 0, 0, 127, 0, 127);
#pragma unroll
    for (int r = 0; r < 16; ++r) d[l * 16 + r] = c[r];
}

__global__ void fp8_convert_kernel(const float *in, unsigned char *out, long long npairs) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npairs;
         i += (long long)gridDim.x * blockDim.x) {
        const int packed = pack_fp8x2(in[2 * i], in[2 * i + 1]);
        out[2 * i] = (unsigned char)(packed & 0xff);
        out[2 * i + 1] = (unsigned char)((packed >> 8) & 0xff);
    }
}

// DPP wave_shl:1 (dpp_ctrl 0x130) of a, with 0 for lanes that have no source.
__global__ void dpp_wave_shl_kernel(const unsigned *a, unsigned *out) {
    const int l = threadIdx.x;
    out[l] = (unsigned)__builtin_amdgcn_update_dpp(0, (int)a[l], 0x130, 0xf, 0xf, false);
}

// DPP whole-wave shift or rotate by one lane (dpp_ctrl CTRL) of a; lanes with
// no source lane keep old[l] (bound_ctrl off), as the advection line kernel
// relies on for its neighbouring wave's edge cells.
template <int CTRL>
__global__ void dpp_wave_move_kernel(const unsigned *a, const unsigned *old, unsigned *out) {
    const int l = threadIdx.x;
    out[l] = (unsigned)__builtin_amdgcn_update_dpp((int)old[l], (int)a[l], CTRL, 0xf, 0xf, false);
}

int done(const char *) { return hipGetLastError() == hipSuccess ? 0 : 1; }

}  // namespace

extern "C" {
// a[64][8], b[64][8] bf16 -> d[64][16] f32
int probe_mfma_bf16(const void *a, const void *b, void *d, void *stream) {
    hipLaunchKernelGGL(mfma_bf16_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, (const __bf16 *)a,
                       (const __bf16 *)b, (float *)d);
    return done("mfma_bf16_kernel");
}
// a[64][32], b[64][32] e4m3 bytes -> d[64][16] f32
int probe_mfma_fp8(const void *a, const void *b, void *d, void *stream) {
    hipLaunchKernelGGL(mfma_fp8_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, (const i32x8_t *)a,
                       (const i32x8_t *)b, (float *)d);
    return done("mfma_fp8_kernel");
}
// a[64] u32 -> out[64]: the lane shift of DPP wave_shl:1
int probe_dpp_wave_shl(const void *a, void *out, void *stream) {
    hipLaunchKernelGGL(dpp_wave_shl_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, (const unsigned *)a,
                       (unsigned *)out);
    return done("dpp_wave_shl_kernel");
}
// a[64], old[64] u32 -> out[64]: which = 0 wave_shl:1, 1 wave_rol:1,
// 2 wave_shr:1, 3 wave_ror:1
int probe_dpp_wave_move(const void *a, const void *old, void *out, unsigned long long which, void *stream) {
    auto *pa = (const unsigned *)a;
    auto *po = (const unsigned *)old;
    auto *pr = (unsigned *)out;
    auto st = (hipStream_t)stream;
    switch (which) {
        case 0: hipLaunchKernelGGL(dpp_wave_move_kernel<0x130>, dim3(1), dim3(64), 0, st, pa, po, pr); break;
        case 1: hipLaunchKernelGGL(dpp_wave_move_kernel<0x134>, dim3(1), dim3(64), 0, st, pa, po, pr); break;
        case 2: hipLaunchKernelGGL(dpp_wave_move_kernel<0x138>, dim3(1), dim3(64), 0, st, pa, po, pr); break;
        case 3: hipLaunchKernelGGL(dpp_wave_move_kernel<0x13C>, dim3(1), dim3(64), 0, st, pa, po, pr); break;
        default: return 1;
    }
    return done("dpp_wave_move_kernel");
}
// a[64] u32 -> out[64]: DPP row_ror:r (r = 1 .. 15) within each 16-lane row
int probe_dpp_row_ror(const void *a, void *out, unsigned long long r, void *stream) {
    auto *pa = (const unsigned *)a;
    auto *pr = (unsigned *)out;
    auto st = (hipStream_t)stream;
#define ROR_CASE(R)                                                                                      \
    case R:                                                                                              \
        hipLaunchKernelGGL(dpp_wave_move_kernel<0x120 + R>, dim3(1), dim3(64), 0, st, pa, pa, pr); break;
    switch (r) {
        ROR_CASE(1) ROR_CASE(2) ROR_CASE(3) ROR_CASE(4) ROR_CASE(5) ROR_CASE(6) ROR_CASE(7) ROR_CASE(8)
        ROR_CASE(9) ROR_CASE(10) ROR_CASE(11) ROR_CASE(12) ROR_CASE(13) ROR_CASE(14) ROR_CASE(15)
        default: return 1;
    }
#undef ROR_CASE
    return done("dpp_wave_move_kernel");
}
// in[n] f32 (n even) -> out[n] e4m3 bytes
int probe_fp8_convert(const void *in, void *out, unsigned long long n, void *stream) {
    if (n % 2) return 1;
    const long long np = (long long)(n / 2);
    const unsigned g = (unsigned)std::min<long long>((np + 255) / 256, 4096);
    hipLaunchKernelGGL(fp8_convert_kernel, dim3(g ? g : 1), dim3(256), 0, (hipStream_t)stream, (const float *)in,
                       (unsigned char *)out, np);
    return done("fp8_convert_kernel");
}
}
