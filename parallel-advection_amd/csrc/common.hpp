// Shared host/device helpers of the gfx950 BKMA library.
#pragma once

#include <type_traits>

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "bkma_amd.h"

namespace bkma {

using i64 = long long;
using u64 = unsigned long long;

// ---------------------------------------------------------------------------
// Error reporting (thread-local last error, status codes instead of throws)
// ---------------------------------------------------------------------------
void set_last_error(const std::string &msg);
bkma_status fail(bkma_status st, const std::string &msg);
bkma_status hip_fail(hipError_t e, const char *what);

#define BKMA_HIP_TRY(call)                                                   \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return ::bkma::hip_fail(_e, #call);            \
    } while (0)

#define BKMA_TRY(call)                                                       \
    do {                                                                     \
        bkma_status _s = (call);                                             \
        if (_s != BKMA_OK) return _s;                                        \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Checks that a gfx950-class device is current and usable; every compute
// entry point calls it so a host without a GPU gets BKMA_ERR_DEVICE.
bkma_status require_device();

// Largest LDS allocation (bytes) one work-group may make on the current device.
bkma_status lds_limit_bytes(u64 *bytes);

// Launch-error check after a <<<>>> launch.
bkma_status check_launch(const char *what);

inline u64 ceil_div(u64 a, u64 b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Counter-based generator (identical on host and device; the oracle uses the
// same constants so CPU and GPU produce bit-identical U[0,1) streams).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ u64 rng_key(u64 seed) { return mix64(seed ^ 0x5851F42D4C957F2DULL); }
__host__ __device__ __forceinline__ double rng_uniform(u64 key, u64 idx) {
    const u64 z = mix64(key + (idx + 1ULL) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 0x1.0p-53;
}

// ---------------------------------------------------------------------------
// Storage conversions
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ double to_f64(T v) { return (double)v; }
template <class T> __device__ __forceinline__ T from_f64(double v) { return (T)v; }

// ---------------------------------------------------------------------------
// Advection parameters as the kernels see them (subset of bkma_adv_params).
// ---------------------------------------------------------------------------
// Compile-time loop: f(integral_constant<int, I>) for I = B .. E-1.
template <int B, int E, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int clcm(int a, int b) { return a / cgcd(a, b) * b; }
constexpr int cmod(int a, int m) { return ((a % m) + m) % m; }

struct AdvDev {
    i64 n0, n1, n2;
    double dt, dx, dvx, inv_dx;
    double minRealX, realWidthX, minRealVx;
};

inline AdvDev make_adv_dev(const bkma_adv_params &p) {
    AdvDev d;
    d.n0 = (i64)p.n0;
    d.n1 = (i64)p.n1;
    d.n2 = (i64)p.n2;
    d.dt = p.dt;
    d.dx = p.dx;
    d.dvx = p.dvx;
    d.inv_dx = p.inv_dx;
    d.minRealX = p.minRealX;
    d.realWidthX = p.realWidthX;
    d.minRealVx = p.minRealVx;
    return d;
}

// The whole library is compiled with -ffp-contract=off: every expression
// below rounds each multiply and add separately, in the reference's order.
// Fused operations are written explicitly (fma) where a kernel wants them.

// coord, src/solvers/AdvectionSolver.hpp:23-26
__device__ __forceinline__ double adv_coord(int i, double minValue, double delta) { return minValue + i * delta; }

// fmod(a, W) for W > 0.  IEEE fmod is exact; for 0 <= a < 2W it equals
// a (a < W) or a - W (exact by Sterbenz since W <= a < 2W).  Other values
// take the library fmod.
__device__ __forceinline__ double periodic_fmod(double a, double W) {
    if (a >= 0.0 && a < W) return a;
    if (a >= W && a < 2.0 * W) return a - W;
    return fmod(a, W);
}

// displ + leftNode + d_prev1, src/solvers/AdvectionSolver.hpp:56-85.
__device__ __forceinline__ int adv_foot(const AdvDev &p, int i0, int i1, double &d_prev1) {
    const double x = adv_coord(i1, p.minRealX, p.dx);
    const double vx = adv_coord(i0, p.minRealVx, p.dvx);
    const double displx = p.dt * vx;
    const double xFoot = p.minRealX + periodic_fmod(p.realWidthX + x - displx - p.minRealX, p.realWidthX);
    const int leftNode = (int)floor((xFoot - p.minRealX) * p.inv_dx);
    d_prev1 = 2 + p.inv_dx * (xFoot - adv_coord(leftNode, p.minRealX, p.dx));
    return leftNode;
}

// lag_basis, src/solvers/AdvectionSolver.hpp:31-51 (products left to right).
__device__ __forceinline__ void adv_lag_basis(double px, double c[6]) {
    const double l0 = -1. / 24, l1 = 1. / 24., l2 = -1. / 12., l3 = 1. / 12., l4 = -1. / 24., l5 = 1. / 24.;
    const double pxm2 = px - 2.;
    const double sqrpxm2 = pxm2 * pxm2;
    const double pxm2_01 = pxm2 * (pxm2 - 1.);
    c[0] = l0 * pxm2_01 * (pxm2 + 1.) * (pxm2 - 2.) * (pxm2 - 1.);
    c[1] = l1 * pxm2_01 * (pxm2 - 2.) * (5 * sqrpxm2 + pxm2 - 8.);
    c[2] = l2 * (pxm2 - 1.) * (pxm2 - 2.) * (pxm2 + 1.) * (5 * sqrpxm2 - 3 * pxm2 - 6.);
    c[3] = l3 * pxm2 * (pxm2 + 1.) * (pxm2 - 2.) * (5 * sqrpxm2 - 7 * pxm2 - 4.);
    c[4] = l4 * pxm2_01 * (pxm2 + 1.) * (5 * sqrpxm2 - 11 * pxm2 - 2.);
    c[5] = l5 * pxm2_01 * pxm2 * (pxm2 + 1.) * (pxm2 - 2.);
}

// First stencil index (n1 + ipos1) % n1 in the reference's unsigned
// arithmetic (AdvectionSolver.hpp:92-93); tap k is then (base + k) mod n1.
// Valid whenever n1 + ipos1 >= 0, which every foot in [minX - W, minX + 2W)
// satisfies; stencil_index_slow covers the rest exactly.
__device__ __forceinline__ bool stencil_fast_ok(i64 n1, int ipos1) { return (i64)ipos1 >= -n1; }
// (n1 + ipos1) % n1 for n1 + ipos1 >= 0, with the division only taken when
// the sum is at least 2*n1.
__device__ __forceinline__ i64 stencil_base(i64 n1, int ipos1) {
    const i64 s = n1 + ipos1;
    if (s < n1) return s;
    if (s - n1 < n1) return s - n1;
    return s % n1;
}
__device__ __forceinline__ i64 stencil_index_slow(i64 n1, int ipos1, int k) {
    return (i64)(((u64)n1 + (u64)(i64)ipos1 + (u64)k) % (u64)n1);
}

// AdvectionSolver (src/solvers/AdvectionSolver.hpp:13-99) in FP64.  `line` is
// an executor's line view (bkma/executors.hpp): line(i) is row i's value.
struct AdvSolverDev {
    AdvDev p;
    __host__ __device__ __forceinline__ i64 window() const { return 1; }

    template <class Line>
    __device__ __forceinline__ double operator()(const Line &line, i64 i0, i64 i1, i64 /*i2*/) const {
        double d_prev1;
        const int leftNode = adv_foot(p, (int)i0, (int)i1, d_prev1);
        double coef[6];
        adv_lag_basis(d_prev1, coef);
        const int ipos1 = leftNode - 2;
        double value = 0.;
        const i64 n1 = p.n1;
        if (stencil_fast_ok(n1, ipos1)) {
            i64 idx = stencil_base(n1, ipos1);
#pragma unroll
            for (int k = 0; k <= 5; k++) {
                value += coef[k] * line(idx);
                idx += 1;
                if (idx >= n1) idx -= n1;
            }
        } else {
            for (int k = 0; k <= 5; k++) value += coef[k] * line(stencil_index_slow(n1, ipos1, k));
        }
        return value;
    }
};

// ConvSolver (src/solvers/ConvSolver.hpp:7-45): weight span (K, C_in, C_out),
// reads only the first L elements of the line (channel 0), flipped kernel.
template <class T> struct ConvSolverDev {
    const T *w;
    const T *b;
    i64 K, C, L, n1;
    __host__ __device__ __forceinline__ i64 window() const { return K; }

    template <class Line>
    __device__ __forceinline__ double operator()(const Line &line, i64 /*i0*/, i64 i1, i64 /*i2*/) const {
        const i64 oc = i1 / L;
        const i64 il = i1 - oc * L;
        double sum = to_f64(b[oc]);
        for (i64 ic = 0; ic < C; ++ic) {
            for (i64 k = 0; k < K; ++k) {
                const int input_idx = (int)(il - k);
                if ((u64)(i64)input_idx < (u64)n1) sum += line(input_idx) * to_f64(w[(k * C + ic) * C + oc]);
            }
        }
        return sum;
    }
};

// The reference's AdaptiveWg executor (FP64 storage) over slabs
// [i0_begin, i0_begin + count) of the [n0][n1][n2] array at data: `o` is the
// dispatch created for `count` slabs, its mem_space chooses LDS or the
// caller's global scratch.  Used by the tuned plans' local/global splits.
bkma_status adaptivewg_slabs(double *data, const bkma_adv_params &p, u64 i0_begin, u64 count,
                             const bkma_optim_params &o, const bkma_span3d &scratch, hipStream_t stream);

}  // namespace bkma
