// Tuned in-place advection for gfx950.
//
// Strategy COLUMN_STREAM (lines strided by n2 >= 64): one wavefront owns
// 64*VEC consecutive i2 columns of one i0 slab and walks the whole line in
// increasing i1.  Each lane keeps a ring of R old rows in registers:
// loads run AHEAD+P rows in front of the row being written, so every old
// value a stencil needs is read from HBM before its row is overwritten --
// the update is in place with exactly one read and one write per cell.
// The periodic wrap needs the first AHEAD old rows again at the end of the
// line; they are kept in LDS.  Tail rows (n1-BACK .. n1-1) are read first,
// before anything is written.
//
// The FP64 foot location (displ, floor, d_prev1) is computed per row, in
// the reference's operation order, by every lane of the wave (all lanes of
// a wave share i0 and i1).  lag_basis is re-evaluated only when d_prev1's
// bits change, which keeps the coefficients bit-identical to a per-cell
// evaluation.  FP64 storage accumulates with separately rounded multiply
// and add as the reference does; FP32 storage accumulates in FP64 with
// fused multiply-adds and rounds once to FP32 at the store.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <type_traits>
#include "common.hpp"

namespace bkma {

template <class T, int VEC> struct VecOf;
template <> struct VecOf<float, 1> { using type = float; };
template <> struct VecOf<float, 2> { using type = float2; };
template <> struct VecOf<float, 4> { using type = float4; };
template <> struct VecOf<double, 1> { using type = double; };
template <> struct VecOf<double, 2> { using type = double2; };

template <class T, int VEC> struct VecIO {
    using V = typename VecOf<T, VEC>::type;
    __device__ __forceinline__ static void load(const T *p, T (&out)[VEC]) {
        const V v = *reinterpret_cast<const V *>(p);
        __builtin_memcpy(out, &v, sizeof(V));
    }
    __device__ __forceinline__ static void store(T *p, const T (&in)[VEC]) {
        V v;
        __builtin_memcpy(&v, in, sizeof(V));
        *reinterpret_cast<V *>(p) = v;
    }
};

// Interpolation of VEC cells from the 6 ring slots s0..s0+5 (mod R).
template <class T, int VEC, int R, int S0>
__device__ __forceinline__ void interp_slots(const T (&win)[R][VEC], const double (&c)[6], T (&out)[VEC]) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        double acc = 0.;
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc = acc + c[k] * (double)win[(S0 + k) % R][v];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc = fma(c[k], (double)win[(S0 + k) % R][v], acc);
        }
        out[v] = (T)acc;
    }
}

// Dispatch on the row's backward shift d = i1 - leftNode in [DLO, DHI]:
// stencil rows are i1 - d - 2 + k, whose ring slot for loop position J is
// (J + BACK - d - 2 + k) mod R.
template <class T, int VEC, int R, int J, int BACK, int D, int DHI>
__device__ __forceinline__ void interp_switch(int d, const T (&win)[R][VEC], const double (&c)[6], T (&out)[VEC]) {
    if constexpr (D > DHI) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) out[v] = (T)NAN;  // unreachable by construction
    } else {
        if (d == D) {
            constexpr int S0 = ((J + BACK - D - 2) % R + R) % R;
            interp_slots<T, VEC, R, S0>(win, c, out);
        } else {
            interp_switch<T, VEC, R, J, BACK, D + 1, DHI>(d, win, c, out);
        }
    }
}

template <int B, int E, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <class T, int VEC, int SMAX, int P>
struct StreamCfg {
    static constexpr int DLO = -(SMAX + 1), DHI = SMAX + 1;
    static constexpr int BACK = DHI + 2;  // rows behind the output row
    static constexpr int AHEAD = 3 - DLO; // rows in front of it
    static constexpr int WS = BACK + AHEAD + 1;
    static constexpr int R = WS + P;      // ring rows
};

struct StreamArgs {
    AdvDev p;
    i64 tiles_per_slab;  // column tiles of 64*VEC per i0 slab
    i64 total_waves;     // n0 * tiles_per_slab
};

template <class T, int VEC, int SMAX, int P, int WPB>
__global__ __launch_bounds__(64 * WPB) void adv_column_stream_kernel(T *__restrict__ data, StreamArgs a) {
    using C = StreamCfg<T, VEC, SMAX, P>;
    constexpr int R = C::R, BACK = C::BACK, AHEAD = C::AHEAD;
    __shared__ T head_lds[WPB][AHEAD][64 * VEC];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const i64 wave = (i64)blockIdx.x * WPB + wv;
    if (wave >= a.total_waves) return;  // whole wave exits together
    const AdvDev &p = a.p;
    const i64 i0 = wave / a.tiles_per_slab;
    const i64 tile = wave - i0 * a.tiles_per_slab;
    const i64 col = (tile * 64 + lane) * VEC;
    const bool lane_ok = col < p.n2;  // VEC divides n2, so the whole vector is valid
    const i64 n1 = p.n1, n2 = p.n2;
    T *slab = data + (i0 * n1) * n2 + (lane_ok ? col : 0);

    T win[R][VEC];
    auto load_row = [&](i64 row, T (&dst)[VEC]) __attribute__((always_inline)) {
        if (lane_ok) {
            VecIO<T, VEC>::load(slab + row * n2, dst);
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) dst[v] = (T)0;
        }
    };

    // Prologue: logical rows q in [-BACK, AHEAD + P) -> slot q + BACK.
    static_for<0, BACK + AHEAD + P>([&](auto sc) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value;
        constexpr int q = S - BACK;
        const i64 row = q < 0 ? n1 + q : q;
        load_row(row, win[S]);
        if constexpr (q >= 0 && q < AHEAD) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) head_lds[wv][q][lane * VEC + v] = win[S][v];
        }
    });

    double coef[6];
    double cached_p = NAN;
    const int half = (int)(n1 / 2);
    i64 base_row = 0;

    // One output row at ring position J (compile-time, so every ring slot
    // below is a constant and the ring lives in registers).
    auto row_step = [&](auto jc) __attribute__((always_inline)) {
        constexpr int J = decltype(jc)::value;
        const i64 q = base_row + J;
        if (q >= n1) return;
        // Prefetch logical row q + AHEAD + P into the slot of row
        // q - BACK - 1, which no later output reads.
        constexpr int slot = (J + R - 1) % R;
        const i64 qn = q + AHEAD + P;
        if (qn < n1) {
            load_row(qn, win[slot]);
        } else if (qn < n1 + AHEAD) {
            const int h = (int)(qn - n1);
#pragma unroll
            for (int hh = 0; hh < AHEAD; ++hh)
                if (h == hh) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v) win[slot][v] = head_lds[wv][hh][lane * VEC + v];
                }
        }
        // FP64 foot location of row q (shared by every lane of the wave).
        double d_prev1;
        const int leftNode = adv_foot(p, (int)i0, (int)q, d_prev1);
        if (!(d_prev1 == cached_p)) {
            adv_lag_basis(d_prev1, coef);
            cached_p = d_prev1;
        }
        int d = (int)q - leftNode;
        if (d > half) d -= (int)n1;
        if (d < -half) d += (int)n1;
        T out[VEC];
        interp_switch<T, VEC, R, J, BACK, C::DLO, C::DHI>(d, win, coef, out);
        if (lane_ok) VecIO<T, VEC>::store(slab + q * n2, out);
    };

    for (; base_row < n1; base_row += R) static_for<0, R>(row_step);
}

// ---------------------------------------------------------------------------
// Host side: variant table, plan, autotuner
// ---------------------------------------------------------------------------
namespace {

using LaunchFn = void (*)(void *data, const StreamArgs &a, unsigned grid, hipStream_t s);

template <class T, int VEC, int SMAX, int P, int WPB>
void launch_stream(void *data, const StreamArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((adv_column_stream_kernel<T, VEC, SMAX, P, WPB>), dim3(grid), dim3(64 * WPB), 0, s,
                       reinterpret_cast<T *>(data), a);
}

struct StreamVariant {
    bkma_dtype dtype;
    int vec, smax, prefetch, wpb;
    LaunchFn fn;
};

#define SV(T, DT, V, S, P, W) StreamVariant{DT, V, S, P, W, &launch_stream<T, V, S, P, W>}
const StreamVariant kStreamVariants[] = {
#ifdef BKMA_TUNED_FEW
    SV(float, BKMA_F32, 4, 2, 4, 1),
#else
    SV(float, BKMA_F32, 4, 2, 4, 1), SV(float, BKMA_F32, 4, 2, 4, 2), SV(float, BKMA_F32, 4, 2, 8, 1),
    SV(float, BKMA_F32, 2, 2, 4, 1), SV(float, BKMA_F32, 2, 2, 8, 1), SV(float, BKMA_F32, 1, 2, 8, 1),
    SV(float, BKMA_F32, 4, 1, 4, 1), SV(float, BKMA_F32, 2, 1, 8, 1), SV(float, BKMA_F32, 1, 1, 8, 1),
    SV(float, BKMA_F32, 4, 4, 4, 1), SV(float, BKMA_F32, 2, 4, 4, 1), SV(float, BKMA_F32, 1, 4, 4, 1),
    SV(double, BKMA_F64, 2, 2, 4, 1), SV(double, BKMA_F64, 2, 2, 8, 1), SV(double, BKMA_F64, 1, 2, 8, 1),
    SV(double, BKMA_F64, 2, 1, 4, 1), SV(double, BKMA_F64, 1, 1, 8, 1),
    SV(double, BKMA_F64, 2, 4, 4, 1), SV(double, BKMA_F64, 1, 4, 4, 1),
#endif
};
#undef SV

// Largest |v * dt / dx| over the velocity grid, with the reference's
// coordinate formulas; the kernel's shift window must cover ceil of it.
double max_abs_shift(const bkma_adv_params &p) {
    const double v0 = p.minRealVx;
    const double v1 = p.minRealVx + (double)(p.n0 - 1) * p.dvx;
    return std::max(std::fabs(v0), std::fabs(v1)) * std::fabs(p.dt) * p.inv_dx;
}

}  // namespace

struct TunedPlan {
    bkma_adv_params params;
    bkma_dtype dtype;
    bkma_tuned_config cfg;
    const StreamVariant *variant = nullptr;
    std::string name;
};

static bkma_status stream_feasible(const bkma_adv_params &p, bkma_dtype dt, const StreamVariant &v,
                                   std::string *why) {
    if (v.dtype != dt) return BKMA_ERR_UNSUPPORTED;
    if (p.n2 % (u64)v.vec != 0) {
        if (why) *why = "n2 not a multiple of VEC";
        return BKMA_ERR_UNSUPPORTED;
    }
    const double s = max_abs_shift(p);
    if (!(s <= (double)v.smax)) {
        if (why) *why = "shift larger than the register window";
        return BKMA_ERR_UNSUPPORTED;
    }
    const int R = (2 * (v.smax + 1) + 6) + v.prefetch;
    if (p.n1 < (u64)(2 * R + 8)) {
        if (why) *why = "line shorter than the register ring";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (p.n1 > 0x7fffffffULL || p.n0 > 0x7fffffffULL) {
        if (why) *why = "n0/n1 beyond int range";
        return BKMA_ERR_UNSUPPORTED;
    }
    return BKMA_OK;
}

static bkma_status plan_launch(const TunedPlan &pl, void *data, hipStream_t s) {
    const StreamVariant &v = *pl.variant;
    StreamArgs a;
    a.p = make_adv_dev(pl.params);
    a.tiles_per_slab = (i64)ceil_div(pl.params.n2, (u64)(64 * v.vec));
    a.total_waves = (i64)pl.params.n0 * a.tiles_per_slab;
    const u64 grid = ceil_div((u64)a.total_waves, (u64)v.wpb);
    if (grid == 0 || grid > 0x7fffffffULL) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: grid out of range");
    v.fn(data, a, (unsigned)grid, s);
    return check_launch("adv_column_stream_kernel");
}

static std::string variant_name(const StreamVariant &v) {
    return std::string("column_stream<") + (v.dtype == BKMA_F64 ? "f64" : "f32") + ",vec" + std::to_string(v.vec) +
           ",smax" + std::to_string(v.smax) + ",p" + std::to_string(v.prefetch) + ",wpb" + std::to_string(v.wpb) + ">";
}

// Time one variant on a scratch copy of (up to) `sub_n0` slabs.
static bkma_status time_variant(const bkma_adv_params &p, const StreamVariant &v, void *buf, int reps,
                                hipStream_t s, float *ms) {
    TunedPlan pl;
    pl.params = p;
    pl.dtype = v.dtype;
    pl.variant = &v;
    hipEvent_t e0, e1;
    BKMA_HIP_TRY(hipEventCreate(&e0));
    BKMA_HIP_TRY(hipEventCreate(&e1));
    bkma_status st = plan_launch(pl, buf, s);  // warm-up
    if (st == BKMA_OK) {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && st == BKMA_OK; ++r) st = plan_launch(pl, buf, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(ms, e0, e1);
        *ms /= (float)reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

}  // namespace bkma

using namespace bkma;

struct bkma_adv_plan : TunedPlan {};

extern "C" bkma_status bkma_adv_plan_create(const bkma_adv_params *params, bkma_dtype dtype, int32_t autotune,
                                            void *stream, bkma_adv_plan **out) {
    if (!params || !out) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (dtype != BKMA_F32 && dtype != BKMA_F64) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: F32 or F64");
    if (params->n0 == 0 || params->n1 == 0 || params->n2 == 0)
        return fail(BKMA_ERR_INVALID_ARGUMENT, "array extents must be positive");
    BKMA_TRY(require_device());

    std::vector<const StreamVariant *> cands;
    std::string why;
    for (const auto &v : kStreamVariants)
        if (stream_feasible(*params, dtype, v, &why) == BKMA_OK) cands.push_back(&v);
    if (cands.empty())
        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: no strategy covers this shape (" + why + ")");

    // Heuristic default: widest vector, smallest window, P = 4, 1 wave/WG.
    auto score = [](const StreamVariant *v) { return v->vec * 1000 - v->smax * 10 - (v->prefetch == 4 ? 0 : 1) - v->wpb; };
    const StreamVariant *best = *std::max_element(cands.begin(), cands.end(),
                                                  [&](auto *a, auto *b) { return score(a) < score(b); });

    if (autotune && cands.size() > 1) {
        // Time candidates on a scratch array holding a sub-problem of whole
        // slabs large enough to fill the GPU several times over.
        bkma_adv_params sub = *params;
        const u64 elems_per_slab = params->n1 * params->n2;
        const u64 target = 256ull << 20;  // ~256 Mi cells
        sub.n0 = std::max<u64>(1, std::min<u64>(params->n0, target / std::max<u64>(elems_per_slab, 1)));
        const size_t bytes = (size_t)(sub.n0 * elems_per_slab) * (dtype == BKMA_F64 ? 8 : 4);
        void *buf = nullptr;
        BKMA_HIP_TRY(hipMalloc(&buf, bytes));
        (void)hipMemsetAsync(buf, 0, bytes, as_stream(stream));
        float best_ms = 1e30f;
        for (auto *v : cands) {
            if (v->smax > 2 * std::max(1, (int)std::ceil(max_abs_shift(*params))) && v->smax > 1) continue;
            float ms = 0;
            if (time_variant(sub, *v, buf, 3, as_stream(stream), &ms) == BKMA_OK && ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        (void)hipFree(buf);
    }

    auto *pl = new (std::nothrow) bkma_adv_plan();
    if (!pl) return fail(BKMA_ERR_OUT_OF_MEMORY, "plan allocation");
    pl->params = *params;
    pl->dtype = dtype;
    pl->variant = best;
    std::memset(&pl->cfg, 0, sizeof(pl->cfg));
    pl->cfg.strategy = BKMA_STRATEGY_COLUMN_STREAM;
    pl->cfg.vec = best->vec;
    pl->cfg.smax = best->smax;
    pl->cfg.prefetch = best->prefetch;
    pl->cfg.waves_per_wg = best->wpb;
    pl->name = variant_name(*best);
    *out = pl;
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_plan_create_with(const bkma_adv_params *params, bkma_dtype dtype,
                                                 const bkma_tuned_config *cfg, bkma_adv_plan **out) {
    if (!params || !cfg || !out) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    BKMA_TRY(require_device());
    if (cfg->strategy != BKMA_STRATEGY_COLUMN_STREAM)
        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: unknown strategy");
    for (const auto &v : kStreamVariants) {
        if (v.dtype == dtype && v.vec == cfg->vec && v.smax == cfg->smax && v.prefetch == cfg->prefetch &&
            v.wpb == cfg->waves_per_wg) {
            std::string why;
            if (stream_feasible(*params, dtype, v, &why) != BKMA_OK)
                return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
            auto *pl = new (std::nothrow) bkma_adv_plan();
            if (!pl) return fail(BKMA_ERR_OUT_OF_MEMORY, "plan allocation");
            pl->params = *params;
            pl->dtype = dtype;
            pl->variant = &v;
            pl->cfg = *cfg;
            pl->name = variant_name(v);
            *out = pl;
            return BKMA_OK;
        }
    }
    return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: no compiled variant matches the configuration");
}

extern "C" bkma_status bkma_adv_plan_describe(const bkma_adv_plan *plan, bkma_tuned_config *cfg, char *name,
                                              uint64_t name_len) {
    if (!plan) return fail(BKMA_ERR_INVALID_ARGUMENT, "null plan");
    if (cfg) *cfg = plan->cfg;
    if (name && name_len) {
        std::strncpy(name, plan->name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_plan_run(bkma_adv_plan *plan, void *data, uint64_t nsteps, void *stream) {
    if (!plan || !data) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    if (reinterpret_cast<uintptr_t>(data) % 16 != 0)
        return fail(BKMA_ERR_INVALID_ARGUMENT, "tuned advection: data must be 16-byte aligned");
    for (uint64_t t = 0; t < nsteps; ++t) BKMA_TRY(plan_launch(*plan, data, as_stream(stream)));
    return BKMA_OK;
}

extern "C" void bkma_adv_plan_destroy(bkma_adv_plan *plan) { delete plan; }
