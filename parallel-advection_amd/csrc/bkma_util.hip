// Library plumbing (errors, device queries), the dispatch policy of the
// reference (create_optim_params and friends), and the harness kernels:
// fills, the analytic L1 validator and error norms.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace bkma {

// ---------------------------------------------------------------------------
// Errors and device checks
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

bkma_status fail(bkma_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

bkma_status hip_fail(hipError_t e, const char *what) {
    return fail(BKMA_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

bkma_status check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return BKMA_OK;
}

bkma_status require_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(BKMA_ERR_DEVICE, "no HIP device available (this library has no CPU fallback)");
    return BKMA_OK;
}

bkma_status lds_limit_bytes(u64 *bytes) {
    int dev = 0;
    BKMA_HIP_TRY(hipGetDevice(&dev));
    static thread_local int cached_dev = -1;
    static thread_local u64 cached = 0;
    if (dev != cached_dev) {
        int v = 0;
        BKMA_HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cached = (u64)v;
        cached_dev = dev;
    }
    *bytes = cached;
    return BKMA_OK;
}

// ---------------------------------------------------------------------------
// Harness kernels
// ---------------------------------------------------------------------------
template <class T>
__global__ void fill_uniform_kernel(T *out, i64 n, u64 key) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        out[i] = (T)rng_uniform(key, (u64)i);
}

template <class T>
__global__ void fill_normal_kernel(T *out, i64 n, u64 key, double stddev) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const double u1 = 1.0 - rng_uniform(key, 2 * (u64)i);  // (0, 1]
        const double u2 = rng_uniform(key, 2 * (u64)i + 1);
        const double r = sqrt(-2.0 * log(u1));
        out[i] = (T)(stddev * r * cos(6.283185307179586 * u2));
    }
}

template <> __global__ void fill_uniform_kernel<__hip_bfloat16>(__hip_bfloat16 *out, i64 n, u64 key) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        out[i] = __float2bfloat16((float)rng_uniform(key, (u64)i));
}

template <> __global__ void fill_normal_kernel<__hip_bfloat16>(__hip_bfloat16 *out, i64 n, u64 key, double stddev) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const double u1 = 1.0 - rng_uniform(key, 2 * (u64)i);
        const double u2 = rng_uniform(key, 2 * (u64)i + 1);
        const double r = sqrt(-2.0 * log(u1));
        out[i] = __float2bfloat16((float)(stddev * r * cos(6.283185307179586 * u2)));
    }
}

template <class T>
__global__ void fill_constant_kernel(T *out, i64 n, double value) {
    const T v = (T)value;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) out[i] = v;
}
template <> __global__ void fill_constant_kernel<__hip_bfloat16>(__hip_bfloat16 *out, i64 n, double value) {
    const __hip_bfloat16 v = __float2bfloat16((float)value);
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) out[i] = v;
}

// fill_buffer_adv (src/tools/init.hpp:31-46)
template <class T>
__global__ void fill_sin4_kernel(T *data, AdvDev p) {
    const i64 total = p.n0 * p.n1 * p.n2;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const i64 i1 = (idx / p.n2) % p.n1;
        const double x = p.minRealX + i1 * p.dx;
        data[idx] = (T)sin(4 * x * M_PI);
    }
}

template <class T> __device__ __forceinline__ double load_f64(const T *p, i64 i) { return (double)p[i]; }
template <> __device__ __forceinline__ double load_f64<__hip_bfloat16>(const __hip_bfloat16 *p, i64 i) {
    return (double)__bfloat162float(p[i]);
}

constexpr int kRedThreads = 256;

// Per-block partials of the analytic L1 error: block (tile2, chunk) sums
// |f - exact| over rows [chunk*rows_per, ...) of (i0, i1) for 64 i2 values.
template <class T>
__global__ void l1_partial_kernel(const T *data, AdvDev p, double t, i64 rows_per, double *partial) {
    const i64 i2 = (i64)blockIdx.x * 64 + (threadIdx.x & 63);
    const i64 sub = threadIdx.x >> 6;  // 4 row-interleaved sub-groups
    const i64 chunk = blockIdx.y;
    const i64 rows = p.n0 * p.n1;
    const i64 r_begin = chunk * rows_per, r_end = min(rows, r_begin + rows_per);
    double s = 0.0;
    if (i2 < p.n2) {
        for (i64 r = r_begin + sub; r < r_end; r += 4) {
            const i64 i0 = r / p.n1, i1 = r % p.n1;
            const double f = load_f64(data, r * p.n2 + i2);
            const double x = p.minRealX + i1 * p.dx;
            const double v = p.minRealVx + i0 * p.dvx;
            s += fabs(f - sin(4 * M_PI * (x - v * t)));
        }
    }
    __shared__ double red[kRedThreads];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 64) {
        const double tot = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] + red[threadIdx.x + 192];
        if (i2 < p.n2) partial[chunk * p.n2 + i2] = tot;
    }
}

template <class A, class R>
__global__ void norms_partial_kernel(const A *a, const R *ref, i64 n, double *partial) {
    double mx = 0.0, mr = 0.0, s2 = 0.0, r2 = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const double av = load_f64(a, i), rv = load_f64(ref, i);
        const double d = av - rv;
        mx = fmax(mx, fabs(d));
        mr = fmax(mr, fabs(rv));
        s2 += d * d;
        r2 += rv * rv;
        if (d != d) mx = d;  // propagate NaN
    }
    __shared__ double sm[4][kRedThreads];
    sm[0][threadIdx.x] = mx;
    sm[1][threadIdx.x] = mr;
    sm[2][threadIdx.x] = s2;
    sm[3][threadIdx.x] = r2;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double a0 = sm[0][threadIdx.x], b0 = sm[0][threadIdx.x + w];
            sm[0][threadIdx.x] = (a0 != a0 || b0 != b0) ? NAN : fmax(a0, b0);
            sm[1][threadIdx.x] = fmax(sm[1][threadIdx.x], sm[1][threadIdx.x + w]);
            sm[2][threadIdx.x] += sm[2][threadIdx.x + w];
            sm[3][threadIdx.x] += sm[3][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partial[blockIdx.x * 4 + k] = sm[k][0];
}

template <class T>
__global__ void sum_region_kernel(const T *data, i64 n0, i64 n1, i64 n2, i64 nw, double *partial) {
    const i64 total = n0 * nw * n2;
    double s = 0.0;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const i64 i2 = idx % n2;
        const i64 i1 = (idx / n2) % nw;
        const i64 i0 = idx / (n2 * nw);
        s += load_f64(data, (i0 * n1 + i1) * n2 + i2);
    }
    __shared__ double sm[kRedThreads];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

// validate_conv1d (src/tools/validation.hpp:114-156): every value at a given
// i1 < nw must be equal across i0 and i2.
template <class T>
__global__ void conv_equal_kernel(const T *data, i64 n0, i64 n1, i64 n2, i64 nw, int *flag) {
    const i64 total = n0 * nw * n2;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const i64 i2 = idx % n2;
        const i64 i1 = (idx / n2) % nw;
        const i64 i0 = idx / (n2 * nw);
        const double v = load_f64(data, (i0 * n1 + i1) * n2 + i2);
        const double v0 = load_f64(data, i1 * n2);  // data(0, i1, 0)
        if (v != v0) atomicOr(flag, 1);
    }
}

unsigned grid_for(i64 n, unsigned block, unsigned cap = 1u << 16) {
    return (unsigned)std::max<u64>(1, std::min<u64>(ceil_div((u64)std::max<i64>(n, 1), block), cap));
}

template <class F> bkma_status with_dtype(bkma_dtype dt, F &&f) {
    switch (dt) {
        case BKMA_F64: return f((double *)nullptr);
        case BKMA_F32: return f((float *)nullptr);
        case BKMA_BF16: return f((__hip_bfloat16 *)nullptr);
    }
    return fail(BKMA_ERR_UNSUPPORTED, "unknown dtype");
}

// Device scratch owned by one host call.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace bkma

using namespace bkma;

// ---------------------------------------------------------------------------
// C ABI: library / device
// ---------------------------------------------------------------------------
extern "C" int bkma_abi_version(void) { return BKMA_AMD_ABI_VERSION; }

extern "C" const char *bkma_last_error(void) { return g_last_error.c_str(); }

extern "C" bkma_status bkma_query_lds_bytes(int device, uint64_t *bytes) {
    if (!bytes) return fail(BKMA_ERR_INVALID_ARGUMENT, "null output");
    BKMA_TRY(require_device());
    int v = 0;
    BKMA_HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    *bytes = (uint64_t)v;
    return BKMA_OK;
}

// ---------------------------------------------------------------------------
// C ABI: parameters and dispatch policy
// ---------------------------------------------------------------------------
extern "C" void bkma_adv_update_deltas(bkma_adv_params *p) {
    // src/config/AdvectionParams.cpp:96-103
    p->realWidthX = p->maxRealX - p->minRealX;
    p->dx = p->realWidthX / (double)p->n1;
    p->dvx = (p->maxRealVx - p->minRealVx) / (double)p->n0;
    p->inv_dx = 1 / p->dx;
}

extern "C" bkma_batch_config bkma_init_1d_blocking(uint64_t n, uint64_t max_batchs) {
    // src/core/bkma_tools.hpp:138-153 (float division, as written there)
    bkma_batch_config b;
    b.batch_size = n < max_batchs ? n : max_batchs;
    const float div = (float)n / (float)max_batchs;
    const float floor_div = std::floor(div);
    const bool div_is_int = div == floor_div;
    b.n_batch = div_is_int ? (uint64_t)div : (uint64_t)(floor_div + 1);
    b.last_batch_size = div_is_int ? max_batchs : (n % max_batchs);
    return b;
}

extern "C" bkma_status bkma_create_optim_params(uint64_t n0, uint64_t n1, uint64_t n2, uint64_t pref_wg_size,
                                                uint64_t seq_size0, uint64_t seq_size2, uint64_t elem_size,
                                                uint64_t local_mem_bytes, bkma_optim_params *out) {
    if (!out) return fail(BKMA_ERR_INVALID_ARGUMENT, "null output");
    if (n0 == 0 || n1 == 0 || n2 == 0 || pref_wg_size == 0 || elem_size == 0)
        return fail(BKMA_ERR_INVALID_ARGUMENT, "create_optim_params: sizes must be positive");
    if (local_mem_bytes == 0) {
        BKMA_TRY(require_device());
        u64 v = 0;
        BKMA_TRY(lds_limit_bytes(&v));
        local_mem_bytes = v;
    }
    // WorkItemDispatch::set_ideal_sizes (src/core/bkma_tools.hpp:74-91)
    uint64_t w0 = 1, w1, w2;
    if (n2 >= pref_wg_size) {
        w1 = 1;
        w2 = pref_wg_size;
    } else if (n1 * n2 >= pref_wg_size) {
        w1 = pref_wg_size / n2;
        w2 = n2;
    } else {
        w0 = pref_wg_size / n1 * n2;  // (pref / n1) * n2, as written there
        w1 = n1;
        w2 = n2;
    }
    // WorkItemDispatch::adjust_sizes_mem_limit (src/core/bkma_tools.hpp:94-107)
    const uint64_t max_elems = local_mem_bytes / elem_size;
    int32_t mem_space = BKMA_MEM_LOCAL;
    const uint64_t total_wi = w0 * w1 * w2;
    if (w2 * n1 >= max_elems) {
        w2 = max_elems / n1;
        w0 = 1;
        if (w2 == 0) {
            // The reference divides by zero here (line longer than the LDS);
            // this library keeps one line per work-group and a global scratch.
            w2 = 1;
            mem_space = BKMA_MEM_GLOBAL;
        }
        w1 = total_wi / w2 * w0;
    }
    if (w1 == 0) w1 = 1;
    // WorkGroupDispatch::set_num_work_groups(n0, n2, seq_size0, seq_size2, w0,
    // w2) exactly as create_optim_params calls it (init.hpp:89-91): the
    // seq sizes arrive in the n_batchs slots, s0_ = s2_ = 1.
    if (seq_size0 == 0 || seq_size2 == 0)
        return fail(BKMA_ERR_INVALID_ARGUMENT, "create_optim_params: seq sizes must be positive");
    if (1 > n0 || 1 > n2)
        return fail(BKMA_ERR_INVALID_ARGUMENT,
                    "s0_ > n0 || s2_ > n2. Sequential size cannot be larger than dimension size.");
    if (w0 > n0 || w2 > n2)
        return fail(BKMA_ERR_INVALID_ARGUMENT,
                    "s0_*w0 > n0 || s2_*w2 > n2. A single work-group cannot process more than dimension size.");
    bkma_optim_params o;
    std::memset(&o, 0, sizeof(o));
    o.dispatch_d0 = {1, n0, n0};
    o.dispatch_d2 = {1, n2, n2};
    o.w0 = w0;
    o.w1 = w1;
    o.w2 = w2;
    o.wg_dispatch.s0 = 1;
    o.wg_dispatch.s2 = 1;
    o.wg_dispatch.g0 = n0 / seq_size0 / 1 / w0;
    o.wg_dispatch.g2 = n2 / seq_size2 / 1 / w2;
    o.mem_space = mem_space;
    *out = o;
    return BKMA_OK;
}

extern "C" bkma_status bkma_scratch_extents(bkma_impl impl, const bkma_optim_params *optim, uint64_t n0,
                                            uint64_t n1, uint64_t n2, uint64_t window, uint64_t *e0, uint64_t *e1,
                                            uint64_t *e2) {
    if (!e0 || !e1 || !e2) return fail(BKMA_ERR_INVALID_ARGUMENT, "null output");
    if (window == 0 || window > n1) return fail(BKMA_ERR_INVALID_ARGUMENT, "window must be in [1, n1]");
    if (impl == BKMA_IMPL_BASIC_RANGE) {
        *e0 = n0;
        *e1 = n1;
        *e2 = n2;
        return BKMA_OK;
    }
    if (impl == BKMA_IMPL_ADAPTIVE_WG) {
        if (!optim) return fail(BKMA_ERR_INVALID_ARGUMENT, "AdaptiveWg needs BkmaOptimParams");
        // Largest batch decides the global-id range (AdaptiveWg.hpp:66-76).
        const uint64_t b0 = std::max(optim->dispatch_d0.batch_size, optim->dispatch_d0.last_batch_size);
        const uint64_t b2 = std::max(optim->dispatch_d2.batch_size, optim->dispatch_d2.last_batch_size);
        const uint64_t w0 = std::min<uint64_t>(optim->w0, b0), w2 = std::min<uint64_t>(optim->w2, b2);
        const uint64_t s0 = std::max<uint64_t>(optim->wg_dispatch.s0, 1), s2 = std::max<uint64_t>(optim->wg_dispatch.s2, 1);
        if (w0 == 0 || w2 == 0) return fail(BKMA_ERR_INVALID_ARGUMENT, "work-group sizes must be positive");
        *e0 = (b0 / s0 / w0) * w0;
        *e1 = (b2 / s2 / w2) * w2;
        *e2 = n1 - (window - 1);
        return BKMA_OK;
    }
    *e0 = *e1 = *e2 = 0;  // NDRange uses LDS only
    return BKMA_OK;
}

// ---------------------------------------------------------------------------
// C ABI: harness kernels
// ---------------------------------------------------------------------------
extern "C" bkma_status bkma_fill_uniform(bkma_dtype dtype, void *out, uint64_t n, uint64_t seed, void *stream) {
    if (!out && n) return fail(BKMA_ERR_INVALID_ARGUMENT, "null output");
    BKMA_TRY(require_device());
    const u64 key = rng_key(seed);
    return with_dtype(dtype, [&](auto *tag) -> bkma_status {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(fill_uniform_kernel<T>, dim3(grid_for((i64)n, 256)), dim3(256), 0, as_stream(stream),
                           (T *)out, (i64)n, key);
        return check_launch("fill_uniform_kernel");
    });
}

extern "C" bkma_status bkma_fill_normal(bkma_dtype dtype, void *out, uint64_t n, uint64_t seed, double stddev,
                                        void *stream) {
    if (!out && n) return fail(BKMA_ERR_INVALID_ARGUMENT, "null output");
    BKMA_TRY(require_device());
    const u64 key = rng_key(seed);
    return with_dtype(dtype, [&](auto *tag) -> bkma_status {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(fill_normal_kernel<T>, dim3(grid_for((i64)n, 256)), dim3(256), 0, as_stream(stream),
                           (T *)out, (i64)n, key, stddev);
        return check_launch("fill_normal_kernel");
    });
}

extern "C" bkma_status bkma_fill_constant(bkma_dtype dtype, void *out, uint64_t n, double value, void *stream) {
    if (!out && n) return fail(BKMA_ERR_INVALID_ARGUMENT, "null output");
    BKMA_TRY(require_device());
    return with_dtype(dtype, [&](auto *tag) -> bkma_status {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(fill_constant_kernel<T>, dim3(grid_for((i64)n, 256)), dim3(256), 0, as_stream(stream),
                           (T *)out, (i64)n, value);
        return check_launch("fill_constant_kernel");
    });
}

extern "C" bkma_status bkma_fill_adv_sin4(bkma_dtype dtype, void *data, const bkma_adv_params *params,
                                          void *stream) {
    if (!data || !params) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    BKMA_TRY(require_device());
    const AdvDev p = make_adv_dev(*params);
    const i64 total = p.n0 * p.n1 * p.n2;
    switch (dtype) {
        case BKMA_F64:
            hipLaunchKernelGGL(fill_sin4_kernel<double>, dim3(grid_for(total, 256)), dim3(256), 0,
                               as_stream(stream), (double *)data, p);
            break;
        case BKMA_F32:
            hipLaunchKernelGGL(fill_sin4_kernel<float>, dim3(grid_for(total, 256)), dim3(256), 0,
                               as_stream(stream), (float *)data, p);
            break;
        default:
            return fail(BKMA_ERR_UNSUPPORTED, "fill_adv_sin4: dtype must be F64 or F32");
    }
    return check_launch("fill_sin4_kernel");
}

extern "C" bkma_status bkma_validate_adv_l1(bkma_dtype dtype, const void *data, const bkma_adv_params *params,
                                            double *highest, double *lowest, void *stream) {
    if (!data || !params || !highest) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    BKMA_TRY(require_device());
    const AdvDev p = make_adv_dev(*params);
    const i64 rows = p.n0 * p.n1;
    const i64 chunks = std::min<i64>(std::max<i64>(1, rows / 64), std::max<i64>(1, 4096 / (i64)ceil_div(p.n2, 64)));
    const i64 rows_per = (i64)ceil_div((u64)rows, (u64)chunks);
    DevBuf buf;
    BKMA_HIP_TRY(hipMalloc(&buf.p, sizeof(double) * (size_t)(chunks * p.n2)));
    const double t = (double)params->maxIter * params->dt;
    const dim3 grid((unsigned)ceil_div((u64)p.n2, 64), (unsigned)chunks);
    bkma_status st = with_dtype(dtype, [&](auto *tag) -> bkma_status {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(l1_partial_kernel<T>, grid, dim3(kRedThreads), 0, as_stream(stream), (const T *)data, p,
                           t, rows_per, (double *)buf.p);
        return check_launch("l1_partial_kernel");
    });
    if (st != BKMA_OK) return st;
    std::vector<double> part((size_t)(chunks * p.n2));
    BKMA_HIP_TRY(hipMemcpyAsync(part.data(), buf.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                as_stream(stream)));
    BKMA_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    double hi = -1.0, lo = INFINITY;
    for (i64 i2 = 0; i2 < p.n2; ++i2) {
        double s = 0.0;
        for (i64 c = 0; c < chunks; ++c) s += part[(size_t)(c * p.n2 + i2)];
        const double l1 = s / (double)(p.n1 * p.n0);
        hi = std::max(hi, l1);
        lo = std::min(lo, l1);
    }
    *highest = hi;
    if (lowest) *lowest = lo;
    return BKMA_OK;
}

extern "C" bkma_status bkma_error_norms(bkma_dtype a_dtype, const void *a, bkma_dtype ref_dtype, const void *ref,
                                        uint64_t n, double out[4], void *stream) {
    if (!a || !ref || !out) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    BKMA_TRY(require_device());
    const unsigned blocks = grid_for((i64)n, kRedThreads, 4096);
    DevBuf buf;
    BKMA_HIP_TRY(hipMalloc(&buf.p, sizeof(double) * 4 * blocks));
    bkma_status st = with_dtype(a_dtype, [&](auto *ta) -> bkma_status {
        using A = std::remove_pointer_t<decltype(ta)>;
        return with_dtype(ref_dtype, [&](auto *tr) -> bkma_status {
            using R = std::remove_pointer_t<decltype(tr)>;
            hipLaunchKernelGGL((norms_partial_kernel<A, R>), dim3(blocks), dim3(kRedThreads), 0, as_stream(stream),
                               (const A *)a, (const R *)ref, (i64)n, (double *)buf.p);
            return check_launch("norms_partial_kernel");
        });
    });
    if (st != BKMA_OK) return st;
    std::vector<double> part(4 * (size_t)blocks);
    BKMA_HIP_TRY(hipMemcpyAsync(part.data(), buf.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                as_stream(stream)));
    BKMA_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    double mx = 0, mr = 0, s2 = 0, r2 = 0;
    bool nan = false;
    for (unsigned b = 0; b < blocks; ++b) {
        if (std::isnan(part[4 * b])) nan = true;
        mx = std::max(mx, part[4 * b]);
        mr = std::max(mr, part[4 * b + 1]);
        s2 += part[4 * b + 2];
        r2 += part[4 * b + 3];
    }
    out[0] = nan ? NAN : mx;
    out[1] = mr;
    out[2] = s2;
    out[3] = r2;
    return BKMA_OK;
}

extern "C" bkma_status bkma_conv_sum_normalize(bkma_dtype dtype, const void *data, uint64_t n0, uint64_t n1,
                                               uint64_t n2, uint64_t nw, double *mean, void *stream) {
    if (!data || !mean) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    if (nw > n1) return fail(BKMA_ERR_INVALID_ARGUMENT, "nw larger than n1");
    BKMA_TRY(require_device());
    const i64 total = (i64)(n0 * nw * n2);
    const unsigned blocks = grid_for(total, kRedThreads, 4096);
    DevBuf buf;
    BKMA_HIP_TRY(hipMalloc(&buf.p, sizeof(double) * blocks));
    bkma_status st = with_dtype(dtype, [&](auto *tag) -> bkma_status {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(sum_region_kernel<T>, dim3(blocks), dim3(kRedThreads), 0, as_stream(stream),
                           (const T *)data, (i64)n0, (i64)n1, (i64)n2, (i64)nw, (double *)buf.p);
        return check_launch("sum_region_kernel");
    });
    if (st != BKMA_OK) return st;
    std::vector<double> part(blocks);
    BKMA_HIP_TRY(hipMemcpyAsync(part.data(), buf.p, sizeof(double) * blocks, hipMemcpyDeviceToHost,
                                as_stream(stream)));
    BKMA_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    double s = 0;
    for (double v : part) s += v;
    *mean = total ? s / (double)total : 0.0;
    return BKMA_OK;
}

extern "C" bkma_status bkma_conv_validate(bkma_dtype dtype, const void *data, uint64_t n0, uint64_t n1, uint64_t n2,
                                          uint64_t nw, int32_t *all_equal, void *stream) {
    if (!data || !all_equal) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    if (nw > n1) return fail(BKMA_ERR_INVALID_ARGUMENT, "nw larger than n1");
    BKMA_TRY(require_device());
    DevBuf flag;
    BKMA_HIP_TRY(hipMalloc(&flag.p, sizeof(int)));
    BKMA_HIP_TRY(hipMemsetAsync(flag.p, 0, sizeof(int), as_stream(stream)));
    const i64 total = (i64)(n0 * nw * n2);
    bkma_status st = with_dtype(dtype, [&](auto *tag) -> bkma_status {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(conv_equal_kernel<T>, dim3(grid_for(total, 256)), dim3(256), 0, as_stream(stream),
                           (const T *)data, (i64)n0, (i64)n1, (i64)n2, (i64)nw, (int *)flag.p);
        return check_launch("conv_equal_kernel");
    });
    if (st != BKMA_OK) return st;
    int h = 0;
    BKMA_HIP_TRY(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, as_stream(stream)));
    BKMA_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    *all_equal = h ? 0 : 1;
    return BKMA_OK;
}
