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

#include "adv_stream.hpp"
#include "common.hpp"

namespace bkma {

// ---------------------------------------------------------------------------
// Host side: variant table, plan, autotuner
// ---------------------------------------------------------------------------
namespace {

using LaunchFn = void (*)(void *data, const StreamArgs &a, unsigned grid, hipStream_t s);
using OccFn = int (*)();

template <class T, int VEC, int SMAX, int P, int WPB, bool NT>
void launch_stream(void *data, const StreamArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((adv_column_stream_kernel<T, VEC, SMAX, P, WPB, NT>), dim3(grid), dim3(64 * WPB), 0, s,
                       reinterpret_cast<T *>(data), a);
}

// Resident work-groups per CU for a variant (0 if the query fails).
template <class T, int VEC, int SMAX, int P, int WPB, bool NT> int occupancy_stream() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, adv_column_stream_kernel<T, VEC, SMAX, P, WPB, NT>,
                                                     64 * WPB, 0) != hipSuccess)
        return 0;
    return blocks;
}

struct StreamVariant {
    bkma_dtype dtype;
    int vec, smax, prefetch, wpb, nt;
    LaunchFn fn;
    OccFn occ;
};

#define SV(T, DT, V, S, P, W, NT) \
    StreamVariant{DT, V, S, P, W, NT, &launch_stream<T, V, S, P, W, NT>, &occupancy_stream<T, V, S, P, W, NT>}
// P is WS/2 or WS (WS = 2*SMAX + 6) so the unrolled loop stays one window long.
const StreamVariant kStreamVariants[] = {
#ifdef BKMA_TUNED_FEW
    SV(float, BKMA_F32, 4, 2, 10, 1, false),
#else
    SV(float, BKMA_F32, 4, 2, 5, 1, false),   SV(float, BKMA_F32, 4, 2, 10, 1, false),
    SV(float, BKMA_F32, 4, 2, 10, 1, true),   SV(float, BKMA_F32, 2, 2, 10, 1, false),
    SV(float, BKMA_F32, 2, 2, 10, 1, true),   SV(float, BKMA_F32, 1, 2, 10, 1, false),
    SV(float, BKMA_F32, 4, 1, 8, 1, false),   SV(float, BKMA_F32, 2, 1, 8, 1, false),
    SV(float, BKMA_F32, 4, 4, 7, 1, false),   SV(float, BKMA_F32, 2, 4, 14, 1, false),
    SV(float, BKMA_F32, 1, 4, 14, 1, false),
    SV(double, BKMA_F64, 2, 2, 5, 1, false),  SV(double, BKMA_F64, 2, 2, 10, 1, false),
    SV(double, BKMA_F64, 2, 2, 10, 1, true),  SV(double, BKMA_F64, 1, 2, 10, 1, false),
    SV(double, BKMA_F64, 2, 1, 8, 1, false),  SV(double, BKMA_F64, 2, 4, 7, 1, false),
    SV(double, BKMA_F64, 1, 4, 14, 1, false),
#endif
};
#undef SV

// Range of the backward shift d = i1 - leftNode over the velocity grid.
// leftNode = floor(i1 - s + e) with s = v*dt/dx and |e| << 1 (rounding), so
// d lies in [floor(s_min), floor(s_max) + 1]; a 1e-6 margin absorbs e.
void shift_range(const bkma_adv_params &p, int *dlo, int *dhi) {
    const double v0 = p.minRealVx;
    const double v1 = p.minRealVx + (double)(p.n0 - 1) * p.dvx;
    const double s0 = v0 * p.dt * p.inv_dx, s1 = v1 * p.dt * p.inv_dx;
    const double smin = std::min(s0, s1), smax = std::max(s0, s1);
    *dlo = (int)std::floor(smin - 1e-6);
    *dhi = (int)std::floor(smax + 1e-6) + 1;
}

}  // namespace

struct TunedPlan {
    bkma_adv_params params;
    bkma_dtype dtype;
    bkma_tuned_config cfg;
    const StreamVariant *variant = nullptr;
    i64 persistent_blocks = 0;  // 0: one wave per unit; else grid cap (resident work-groups)
    std::string name;
};

static bkma_status stream_feasible(const bkma_adv_params &p, bkma_dtype dt, const StreamVariant &v,
                                   std::string *why) {
    if (v.dtype != dt) return BKMA_ERR_UNSUPPORTED;
    if (p.n2 % (u64)v.vec != 0) {
        if (why) *why = "n2 not a multiple of VEC";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (!std::isfinite(p.dt) || !std::isfinite(p.minRealVx) || !std::isfinite(p.dvx) || !(p.realWidthX > 0)) {
        if (why) *why = "non-finite parameters";
        return BKMA_ERR_UNSUPPORTED;
    }
    int dlo = 0, dhi = 0;
    if (std::fabs(p.dt * p.inv_dx * std::max(std::fabs(p.minRealVx), std::fabs(p.maxRealVx))) > 1e6) {
        if (why) *why = "shift larger than the register window";
        return BKMA_ERR_UNSUPPORTED;
    }
    shift_range(p, &dlo, &dhi);
    if (dlo < -v.smax || dhi > v.smax) {
        if (why) *why = "shift larger than the register window";
        return BKMA_ERR_UNSUPPORTED;
    }
    const int WS = 2 * v.smax + 6;
    if (p.n1 < (u64)(2 * (WS + v.prefetch) + 8)) {
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
    a.total_units = (i64)pl.params.n0 * a.tiles_per_slab;
    u64 grid = ceil_div((u64)a.total_units, (u64)v.wpb);
    if (pl.persistent_blocks > 0) grid = std::min<u64>(grid, (u64)pl.persistent_blocks);
    grid = std::min<u64>(grid, 0x7fffffffULL / (u64)(64 * v.wpb));
    if (grid == 0) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: empty grid");
    a.wave_stride = (i64)grid * v.wpb;
    v.fn(data, a, (unsigned)grid, s);
    return check_launch("adv_column_stream_kernel");
}

static std::string variant_name(const StreamVariant &v, bool persistent) {
    return std::string("column_stream<") + (v.dtype == BKMA_F64 ? "f64" : "f32") + ",vec" + std::to_string(v.vec) +
           ",smax" + std::to_string(v.smax) + ",p" + std::to_string(v.prefetch) + ",wpb" + std::to_string(v.wpb) +
           (v.nt ? ",nt" : "") + (persistent ? ",persistent" : "") + ">";
}

// Resident work-groups on the whole device for a variant (persistent grid).
static i64 resident_blocks(const StreamVariant &v) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return (i64)v.occ() * cus;
}

// Time one variant on a scratch copy of (up to) `sub_n0` slabs.
static bkma_status time_variant(const bkma_adv_params &p, const StreamVariant &v, i64 persistent_blocks, void *buf,
                                int reps, hipStream_t s, float *ms) {
    TunedPlan pl;
    pl.params = p;
    pl.dtype = v.dtype;
    pl.variant = &v;
    pl.persistent_blocks = persistent_blocks;
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

static bkma_adv_plan *make_plan(const bkma_adv_params &p, bkma_dtype dtype, const StreamVariant &v,
                                i64 persistent_blocks);

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
    int smin_needed = 1 << 30;
    for (const auto &v : kStreamVariants)
        if (stream_feasible(*params, dtype, v, &why) == BKMA_OK) {
            cands.push_back(&v);
            smin_needed = std::min(smin_needed, v.smax);
        }
    if (cands.empty())
        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: no strategy covers this shape (" + why + ")");

    // Heuristic default: smallest window that covers the shift, widest vector,
    // deepest prefetch, temporal stores, one wave per unit.
    auto score = [](const StreamVariant *v) {
        return -v->smax * 100000 + v->vec * 1000 + v->prefetch * 10 - v->nt - v->wpb;
    };
    const StreamVariant *best = *std::max_element(cands.begin(), cands.end(),
                                                  [&](auto *a, auto *b) { return score(a) < score(b); });
    i64 best_persist = 0;

    if (autotune && cands.size() > 1) {
        // Time candidates (both grid modes) on a scratch array of whole
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
            if (v->smax > smin_needed + 2) continue;
            for (int mode = 0; mode < 2; ++mode) {
                const i64 pb = mode ? resident_blocks(*v) : 0;
                if (mode && pb <= 0) continue;
                float ms = 0;
                if (time_variant(sub, *v, pb, buf, 3, as_stream(stream), &ms) == BKMA_OK && ms < best_ms) {
                    best_ms = ms;
                    best = v;
                    best_persist = pb;
                }
            }
        }
        (void)hipFree(buf);
        (void)hipGetLastError();
    }
    bkma_adv_plan *pl = make_plan(*params, dtype, *best, best_persist);
    if (!pl) return fail(BKMA_ERR_OUT_OF_MEMORY, "plan allocation");
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
            v.wpb == cfg->waves_per_wg && v.nt == (cfg->nontemporal != 0)) {
            std::string why;
            if (stream_feasible(*params, dtype, v, &why) != BKMA_OK)
                return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
            const i64 pb = cfg->persistent ? resident_blocks(v) : 0;
            bkma_adv_plan *pl = make_plan(*params, dtype, v, pb);
            if (!pl) return fail(BKMA_ERR_OUT_OF_MEMORY, "plan allocation");
            *out = pl;
            return BKMA_OK;
        }
    }
    return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: no compiled variant matches the configuration");
}

static bkma_adv_plan *make_plan(const bkma_adv_params &p, bkma_dtype dtype, const StreamVariant &v,
                                i64 persistent_blocks) {
    auto *pl = new (std::nothrow) bkma_adv_plan();
    if (!pl) return nullptr;
    pl->params = p;
    pl->dtype = dtype;
    pl->variant = &v;
    pl->persistent_blocks = persistent_blocks;
    std::memset(&pl->cfg, 0, sizeof(pl->cfg));
    pl->cfg.strategy = BKMA_STRATEGY_COLUMN_STREAM;
    pl->cfg.vec = v.vec;
    pl->cfg.smax = v.smax;
    pl->cfg.prefetch = v.prefetch;
    pl->cfg.waves_per_wg = v.wpb;
    pl->cfg.nontemporal = v.nt;
    pl->cfg.persistent = persistent_blocks > 0;
    pl->name = variant_name(v, persistent_blocks > 0);
    return pl;
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

extern "C" int32_t bkma_adv_plan_variants(bkma_dtype dtype, bkma_tuned_config *out, int32_t max_out) {
    int32_t n = 0;
    for (const auto &v : kStreamVariants) {
        if (v.dtype != dtype) continue;
        if (out && n < max_out) {
            bkma_tuned_config c;
            std::memset(&c, 0, sizeof(c));
            c.strategy = BKMA_STRATEGY_COLUMN_STREAM;
            c.vec = v.vec;
            c.smax = v.smax;
            c.prefetch = v.prefetch;
            c.waves_per_wg = v.wpb;
            c.nontemporal = v.nt;
            out[n] = c;
        }
        ++n;
    }
    return n;
}
