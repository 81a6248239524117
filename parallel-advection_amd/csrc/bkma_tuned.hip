// Tuned in-place advection for gfx950: plans, autotuner and C ABI.
//
// Strategy COLUMN_STREAM (lines strided by n2 >= 64): one wavefront owns
// 64*VEC consecutive i2 columns of one i0 slab and walks the whole line in
// increasing i1 with a register window per fused time step (adv_stream.hpp);
// a plan runs `steps_per_pass` steps per pass over HBM.
// Strategy COLUMN_PAIR: the column stream for slabs whose rows take at most
// two adjacent shifts (adv_pair.hpp): seven-tap windows, no shift branches.
// Strategy LINE_LDS (any n2, any shift): slabs or segments of them staged in
// LDS (adv_lds.hpp), up to 4 steps per pass.
// Strategy LINE_SHFL (contiguous lines, n2 = 1, of 256 k <= 4096 cells, or
// of 16): each line held in registers by one work-group (or one 16-lane row)
// for a pass of up to 256 steps, stencil neighbours through DPP wavefront
// shifts or row rotations (adv_shfl.hpp).
//
// FP64 storage: results bit-identical to the reference's FP64 steps.
// FP32 storage: FP64 foot location / index / lag_basis exactly as the
// reference, FP64 fused accumulation, FP64 intermediate steps inside a pass
// and one rounding to FP32 per pass.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <climits>

#include "adv_lds.hpp"
#include "adv_pair.hpp"
#include "adv_shfl.hpp"
#include "adv_stream.hpp"
#include "common.hpp"

namespace bkma {

constexpr int kShiftStats = 5;  // entries of the slab shift statistics

__global__ void slab_shift_scan_kernel(AdvDev p, int *__restrict__ shift, int *__restrict__ stats) {
    __shared__ int smin, smax;
    const int half = (int)(p.n1 / 2);
    for (i64 i0 = blockIdx.x; i0 < p.n0; i0 += gridDim.x) {
        if (threadIdx.x == 0) {
            smin = INT_MAX;
            smax = INT_MIN;
        }
        __syncthreads();
        int lo = INT_MAX, hi = INT_MIN;
        for (i64 row = threadIdx.x; row < p.n1; row += blockDim.x) {
            double d_prev1;
            const int leftNode = adv_foot(p, (int)i0, (int)row, d_prev1);
            int d = (int)row - leftNode;
            if (d > half) d -= (int)p.n1;
            if (d < -half) d += (int)p.n1;
            lo = min(lo, d);
            hi = max(hi, d);
        }
        atomicMin(&smin, lo);
        atomicMax(&smax, hi);
        __syncthreads();
        if (threadIdx.x == 0) {
            shift[2 * i0] = smin;
            shift[2 * i0 + 1] = smax;
            atomicMin(&stats[0], smin);
            atomicMax(&stats[1], smax);
            atomicMax(&stats[2], smax - smin);
            atomicMin(&stats[smax == smin ? 3 : 4], smax);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Host side: variant table, plan, autotuner
// ---------------------------------------------------------------------------
namespace {

// The compiled column-stream and column-pair variants of a storage type.
const std::vector<StreamVariant> &stream_variants(bkma_dtype dt) {
    static const std::vector<StreamVariant> none;
    auto gather = [](const StreamVariant *(*ring)(int *), const StreamVariant *(*pair)(int *)) {
        std::vector<StreamVariant> all;
        int n = 0;
        const StreamVariant *t = ring(&n);
        all.insert(all.end(), t, t + n);
        t = pair(&n);
        all.insert(all.end(), t, t + n);
        return all;
    };
    static const std::vector<StreamVariant> f32 = gather(stream_variants_f32, pair_variants_f32);
    static const std::vector<StreamVariant> f64 = gather(stream_variants_f64, pair_variants_f64);
    return dt == BKMA_F32 ? f32 : dt == BKMA_F64 ? f64 : none;
}

int steps_per_pass(const bkma_tuned_config &c) { return c.steps_per_pass > 1 ? c.steps_per_pass : 1; }

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

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
struct TunedPlan {
    bkma_adv_params params;
    bkma_dtype dtype;
    bkma_tuned_config cfg;
    std::string name;
    // COLUMN_STREAM / COLUMN_PAIR
    const StreamVariant *variant = nullptr;
    int ns = 1;                        // steps per pass
    i64 persistent_blocks[9] = {};    // by steps per pass: 0 = one wave per unit, else the grid cap
    int *slab_shift = nullptr;         // COLUMN_PAIR, LINE_SHFL: smallest and largest shift of each slab, [n0][2] (owned)
    int shfl_waves = 0;                // LINE_SHFL: wavefronts per line
    int row_wmin = 0, row_tw = 0;      // LINE_SHFL, 16-cell lines: window offsets [row_wmin, row_wmin + row_tw)
    bkma_optim_params optim{};         // PLAIN, HYBRID: the reference's dispatch for its AdaptiveWg executor
    // HYBRID: slabs [0, k_local) with their lines' scratch in LDS (optim),
    // [k_local, n0) with a global scratch (optim_global, gscratch), the two
    // launches concurrent on the caller's stream and a side stream
    u64 k_local = 0;
    bkma_optim_params optim_global{};
    bkma_span3d gscratch{nullptr, 0, 0, 0};  // owned
    hipStream_t side = nullptr;              // owned
    hipEvent_t fork = nullptr, join = nullptr;  // owned
    int pair_ahead = 0;                // COLUMN_PAIR: largest stage lag, 3 - the smallest shift
    // LINE_LDS
    LdsArgs lds{};
    u64 lds_bytes = 0;
    int lds_threads = 256;  // threads per work-group: 256, 512 or 1024
    void *halo = nullptr;  // segmented mode: halo snapshot buffer (owned)
    ~TunedPlan() {
        if (halo) (void)hipFree(halo);
        if (slab_shift) (void)hipFree(slab_shift);
        if (gscratch.data) (void)hipFree(gscratch.data);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace {

constexpr int kLdsMaxSteps = 4;  // LINE_LDS: steps fused per pass

bool params_finite(const bkma_adv_params &p) {
    return std::isfinite(p.dt) && std::isfinite(p.minRealVx) && std::isfinite(p.dvx) && std::isfinite(p.inv_dx) &&
           std::isfinite(p.minRealX) && p.realWidthX > 0 && std::isfinite(p.realWidthX);
}

// Every tuned strategy assumes |v*dt/dx| + 4 below n1/2: a stencil offset
// from its row, normalised into (-n1/2, n1/2], is then the true one.
bool shift_supported(const bkma_adv_params &p, std::string *why) {
    if (!params_finite(p)) {
        if (why) *why = "non-finite parameters";
        return false;
    }
    if (p.n1 > 0x7fffffffULL || p.n0 > 0x7fffffffULL) {
        if (why) *why = "n0/n1 beyond int range";
        return false;
    }
    int dlo, dhi;
    const double smax = std::fabs(p.dt * p.inv_dx) * std::max(std::fabs(p.minRealVx), std::fabs(p.maxRealVx));
    if (!(smax + 4 < (double)p.n1 / 2)) {
        if (why) *why = "shift of half a line or more";
        return false;
    }
    shift_range(p, &dlo, &dhi);
    return true;
}

// Shifts of every (i0, i1) from the device's foot arithmetic: stats =
// {smallest shift, largest shift, largest spread within one slab, smallest
// shift of the slabs with one shift, smallest largest shift of the slabs with
// several (INT_MAX where there are none)}; the
// per-slab smallest and largest shifts ([n0][2]) go to *dlo_out (device
// memory, owned by the caller) when it is not null.
bkma_status scan_slab_shifts(const bkma_adv_params &p, int **dlo_out, int stats[kShiftStats]) {
    int *dlo = nullptr, *dst = nullptr;
    const int init[kShiftStats] = {INT_MAX, INT_MIN, 0, INT_MAX, INT_MAX};
    if (hipMalloc(&dlo, std::max<u64>(p.n0, 1) * 2 * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(BKMA_ERR_OUT_OF_MEMORY, "tuned advection: slab shift table allocation failed");
    }
    if (hipMalloc(&dst, sizeof(init)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(dlo);
        return fail(BKMA_ERR_OUT_OF_MEMORY, "tuned advection: slab shift table allocation failed");
    }
    bkma_status st = BKMA_OK;
    if (hipMemcpy(dst, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) st = fail(BKMA_ERR_RUNTIME, "hipMemcpy");
    if (st == BKMA_OK) {
        const unsigned grid = (unsigned)std::min<u64>(p.n0, 65536);
        hipLaunchKernelGGL(slab_shift_scan_kernel, dim3(grid), dim3(256), 0, 0, make_adv_dev(p), dlo, dst);
        st = check_launch("slab_shift_scan_kernel");
    }
    if (st == BKMA_OK && hipMemcpy(stats, dst, sizeof(init), hipMemcpyDeviceToHost) != hipSuccess)
        st = fail(BKMA_ERR_RUNTIME, "hipMemcpy");
    (void)hipFree(dst);
    if (st == BKMA_OK && dlo_out) {
        *dlo_out = dlo;
    } else {
        (void)hipFree(dlo);
    }
    return st;
}

// Column-pair feasibility given the slab shift statistics.
bkma_status pair_feasible(const bkma_adv_params &p, bkma_dtype dt, const StreamVariant &v, int ns,
                          const int stats[kShiftStats], std::string *why) {
    if (v.dtype != dt || !v.pair) return BKMA_ERR_UNSUPPORTED;
    if (ns < 1 || ns > v.nsmax) {
        if (why) *why = "steps per pass not compiled for this variant";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (!shift_supported(p, why)) return BKMA_ERR_UNSUPPORTED;
    if (p.n2 % (u64)v.vec != 0) {
        if (why) *why = "n2 not a multiple of VEC";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (stats[2] > 1) {
        if (why) *why = "a slab takes more than two shifts";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (stats[0] < -v.smax || stats[1] > v.smax) {
        if (why) *why = "shift larger than the compiled pair range";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (p.n1 < (u64)(ns * (2 * v.smax + 6) + v.prefetch + kPairTaps + 8)) {
        if (why) *why = "line shorter than the register ring";
        return BKMA_ERR_UNSUPPORTED;
    }
    return BKMA_OK;
}

bkma_status stream_feasible(const bkma_adv_params &p, bkma_dtype dt, const StreamVariant &v, int ns,
                            std::string *why) {
    if (v.dtype != dt || v.pair) return BKMA_ERR_UNSUPPORTED;
    if (ns < 1 || ns > v.nsmax) {
        if (why) *why = "steps per pass not compiled for this variant";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (!shift_supported(p, why)) return BKMA_ERR_UNSUPPORTED;
    if (p.n2 % (u64)v.vec != 0) {
        if (why) *why = "n2 not a multiple of VEC";
        return BKMA_ERR_UNSUPPORTED;
    }
    int dlo = 0, dhi = 0;
    shift_range(p, &dlo, &dhi);
    if (dlo < -v.smax || dhi > v.smax) {
        if (why) *why = "shift larger than the register window";
        return BKMA_ERR_UNSUPPORTED;
    }
    const int WS = 2 * v.smax + 6;
    if (p.n1 < (u64)(ns * (WS - 1) + v.prefetch + WS + 8)) {
        if (why) *why = "line shorter than the register ring";
        return BKMA_ERR_UNSUPPORTED;
    }
    return BKMA_OK;
}

// LINE_SHFL feasibility for `cells` consecutive cells per lane: contiguous
// lines of 64 * cells * NW cells (NW = 1 .. 8 wavefronts; 1 .. 16 with 4
// cells), every line's rows within two adjacent shifts, and every line's
// window start -dmax - 2 in the range the lanes g-1 .. g+1 cover (stats: the
// slab shift statistics; lines with two shifts need one cell more).
// cells = 1: lines of 16 cells, one per 16-lane DPP row; every cell's window
// covers the offsets of all lines, [*wmin, *wmin + *tw), with *wmin in
// [-6, 0] and *tw at most 8.
bkma_status shfl_feasible(const bkma_adv_params &p, int cells, int ns, const int stats[kShiftStats], int *nw,
                          std::string *why, int *wmin = nullptr, int *tw = nullptr) {
    if (!shift_supported(p, why)) return BKMA_ERR_UNSUPPORTED;
    if (p.n2 != 1) {
        if (why) *why = "LINE_SHFL needs contiguous lines (n2 = 1)";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (ns < 1 || ns > kShflMaxSteps) {
        if (why) *why = "LINE_SHFL fuses 1 to 256 steps per pass";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (cells == 1) {
        if (p.n1 != 16) {
            if (why) *why = "LINE_SHFL with one cell per lane needs lines of 16 cells";
            return BKMA_ERR_UNSUPPORTED;
        }
        if (stats[0] > stats[1]) {
            if (why) *why = "no shift statistics";
            return BKMA_ERR_UNSUPPORTED;
        }
        const int rlo = -stats[1] - 2, rhi = -stats[0] - 2;  // first-tap offsets of all cells
        if (rlo < -6 || rlo > 0 || rhi - rlo > 2) {
            if (why) *why = "16-cell lines: shifts beyond the compiled windows";
            return BKMA_ERR_UNSUPPORTED;
        }
        if (wmin) *wmin = rlo;
        if (tw) *tw = rhi - rlo + 6;
        *nw = 4;
        return BKMA_OK;
    }
    if (cells != 4 && cells != 8) {
        if (why) *why = "LINE_SHFL holds 1, 4 or 8 cells per lane";
        return BKMA_ERR_UNSUPPORTED;
    }
    const u64 per_wave = 64 * (u64)cells;
    const u64 w = p.n1 / per_wave;
    if (p.n1 % per_wave != 0 || w < 1 || w > (cells == 4 ? 16u : 8u)) {
        if (why) *why = cells == 4 ? "LINE_SHFL with 4 cells per lane needs n1 = 256 k, k = 1 .. 16"
                                   : "LINE_SHFL with 8 cells per lane needs n1 = 512 k, k = 1 .. 8";
        return BKMA_ERR_UNSUPPORTED;
    }
    if (stats[2] > 1) {
        if (why) *why = "a line takes more than two shifts";
        return BKMA_ERR_UNSUPPORTED;
    }
    // window start -dmax - 2 of every line: at least -cells, at most
    // cells - 5 (six-cell windows, one shift) or cells - 6 (seven, two)
    const bool low_single = stats[3] != INT_MAX && -stats[3] - 2 > cells - 5;
    const bool low_pair = stats[4] != INT_MAX && -stats[4] - 2 > cells - 6;
    if (-stats[1] - 2 < -cells || low_single || low_pair) {
        if (why) *why = "shift beyond the neighbouring lanes";
        return BKMA_ERR_UNSUPPORTED;
    }
    *nw = (int)w;
    return BKMA_OK;
}

// LINE_LDS shape for a given segment length (0 = whole line).
bkma_status lds_shape(const bkma_adv_params &p, bkma_dtype dt, i64 seg_rows, i64 chunk_segs, int ns, LdsArgs *a,
                      u64 *bytes, std::string *why) {
    if (!shift_supported(p, why)) return BKMA_ERR_UNSUPPORTED;
    u64 lds_max = 0;
    BKMA_TRY(lds_limit_bytes(&lds_max));
    const u64 esz = dt == BKMA_F64 ? 8 : 4;
    int dlo, dhi;
    shift_range(p, &dlo, &dhi);
    a->p = make_adv_dev(p);
    if (seg_rows <= 0 || (u64)seg_rows >= p.n1) {
        a->seg_rows = (i64)p.n1;
        a->nseg = 1;
        a->hb_max = a->ha_max = 0;
    } else {
        a->seg_rows = seg_rows;
        a->nseg = (i64)ceil_div(p.n1, (u64)seg_rows);
        a->hb_max = std::max(0, dhi + 2);
        a->ha_max = std::max(0, 3 - dlo);
    }
    a->chunk_segs = a->nseg > 1 ? std::max<i64>(1, std::min<i64>(chunk_segs, a->nseg)) : 1;
    // whole lines: chunk_segs is the number of consecutive slabs a unit
    // stages together (short lines)
    a->slabs = a->nseg == 1 ? std::max<i64>(1, std::min<i64>(chunk_segs, (i64)p.n0)) : 1;
    a->nchunk = (i64)ceil_div((u64)a->nseg, (u64)a->chunk_segs);
    a->ns = std::max(1, ns);
    if (a->chunk_segs > 1 && a->seg_rows < a->ns * a->hb_max) {
        if (why) *why = "segments shorter than the halo cannot be chained";
        return BKMA_ERR_UNSUPPORTED;
    }
    auto need = [&](i64 slabs) {
        // staged rows: the segment (or whole lines) and ns steps' halos
        const u64 rows = (u64)(a->seg_rows + a->ns * (a->hb_max + a->ha_max));
        u64 b = (u64)slabs * rows * p.n2 * esz;
        if (a->ns > 1)  // FP64 work buffers of the fused steps, 16-byte aligned after the staged rows
            b = ((b + 15) & ~(u64)15) + (a->nseg > 1 ? 2 : (u64)(a->ns > 2 ? 2 : 1)) * slabs * rows * p.n2 * 8;
        return b;
    };
    *bytes = need(a->slabs);
    if (*bytes > lds_max) {
        if (why) *why = "segment does not fit in LDS";
        return BKMA_ERR_UNSUPPORTED;
    }
    if ((u64)p.n0 * (u64)a->nseg > 0x7fffffffULL) {
        if (why) *why = "too many work units";
        return BKMA_ERR_UNSUPPORTED;
    }
    return BKMA_OK;
}

template <class T, int NT>
bkma_status launch_lds_nt(const TunedPlan &pl, void *data, hipStream_t s, int ns) {
    LdsArgs a = pl.lds;
    a.ns = ns;
    const i64 nunits = a.slabs > 1 ? ((i64)pl.params.n0 + a.slabs - 1) / a.slabs : (i64)pl.params.n0 * a.nseg;
    const unsigned units = (unsigned)std::min<i64>(nunits, kLdsMaxGrid);
    if (a.nseg > 1) {
        // halo snapshot per chunk: the chunk's rows seen as one long segment
        LdsArgs sa = a;
        sa.seg_rows = a.seg_rows * a.chunk_segs;
        sa.nseg = a.nchunk;
        const unsigned cunits = (unsigned)std::min<i64>((i64)pl.params.n0 * a.nchunk, kLdsMaxGrid);
        hipLaunchKernelGGL(halo_snapshot_kernel<T>, dim3(cunits), dim3(256), 0, s, (const T *)data, (T *)pl.halo, sa);
        BKMA_TRY(check_launch("halo_snapshot_kernel"));
        auto k = ns > 1 ? adv_line_chunk_kernel<T, NT, true> : adv_line_chunk_kernel<T, NT, false>;
        BKMA_HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds_bytes));
        hipLaunchKernelGGL(k, dim3(cunits), dim3(NT), (size_t)pl.lds_bytes, s, (T *)data, (const T *)pl.halo, a);
    } else {
        auto k = adv_line_lds_kernel<T, false, NT>;
        BKMA_HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds_bytes));
        hipLaunchKernelGGL(k, dim3(units), dim3(NT), (size_t)pl.lds_bytes, s, (T *)data, (const T *)nullptr, a);
    }
    return check_launch("adv_line_lds_kernel");
}

// LINE_LDS work-groups of 4, 8 or 16 wavefronts (one work-group per CU when
// the staged rows take most of the LDS: more waves hide the HBM latency)
template <class T>
bkma_status launch_lds(const TunedPlan &pl, void *data, hipStream_t s, int ns) {
    switch (pl.lds_threads) {
        case 512: return launch_lds_nt<T, 512>(pl, data, s, ns);
        case 1024: return launch_lds_nt<T, 1024>(pl, data, s, ns);
        default: return launch_lds_nt<T, 256>(pl, data, s, ns);
    }
}

// One pass of ns steps (ns <= the plan's steps per pass).
bkma_status plan_launch(const TunedPlan &pl, void *data, hipStream_t s, int ns) {
    if (pl.cfg.strategy == BKMA_STRATEGY_PLAIN) {
        const bkma_span3d none{nullptr, 0, 0, 0};
        return bkma_adv_run(BKMA_IMPL_ADAPTIVE_WG, pl.dtype, data, &pl.params, &pl.optim, none, s);
    }
    if (pl.cfg.strategy == BKMA_STRATEGY_HYBRID) {
        const bkma_span3d none{nullptr, 0, 0, 0};
        const u64 n0 = pl.params.n0, kg = n0 - pl.k_local;
        double *d = (double *)data;
        if (pl.k_local == 0) return adaptivewg_slabs(d, pl.params, 0, n0, pl.optim_global, pl.gscratch, s);
        if (kg == 0) return adaptivewg_slabs(d, pl.params, 0, n0, pl.optim, none, s);
        // fork: the global-scratch slabs run on the side stream beside the
        // LDS slabs; join before anything later on the caller's stream
        BKMA_HIP_TRY(hipEventRecord(pl.fork, s));
        BKMA_HIP_TRY(hipStreamWaitEvent(pl.side, pl.fork, 0));
        bkma_status st = adaptivewg_slabs(d, pl.params, pl.k_local, kg, pl.optim_global, pl.gscratch, pl.side);
        if (st == BKMA_OK) st = adaptivewg_slabs(d, pl.params, 0, pl.k_local, pl.optim, none, s);
        BKMA_HIP_TRY(hipEventRecord(pl.join, pl.side));
        BKMA_HIP_TRY(hipStreamWaitEvent(s, pl.join, 0));
        return st;
    }
    if (pl.cfg.strategy == BKMA_STRATEGY_LINE_LDS)
        return pl.dtype == BKMA_F64 ? launch_lds<double>(pl, data, s, ns) : launch_lds<float>(pl, data, s, ns);
    if (pl.cfg.strategy == BKMA_STRATEGY_LINE_SHFL) {
        ShflArgs a;
        a.p = make_adv_dev(pl.params);
        a.slab_shift = pl.slab_shift;
        a.ns = ns;
        if (pl.cfg.vec == 1) return launch_line_row16(pl.dtype, pl.row_wmin, pl.row_tw, data, a, (i64)pl.params.n0, s);
        return launch_line_shfl(pl.dtype, pl.cfg.vec, pl.shfl_waves, data, a, (i64)pl.params.n0, s);
    }
    if (ns < 1 || ns > pl.variant->nsmax) return fail(BKMA_ERR_RUNTIME, "tuned advection: steps per pass out of range");
    const StreamVariant &v = *pl.variant;
    StreamArgs a;
    a.p = make_adv_dev(pl.params);
    a.tiles_per_slab = (i64)ceil_div(pl.params.n2, (u64)(64 * v.vec));
    a.total_units = (i64)pl.params.n0 * a.tiles_per_slab;
    u64 grid = ceil_div((u64)a.total_units, (u64)v.wpb);
    if (pl.persistent_blocks[ns] > 0) grid = std::min<u64>(grid, (u64)pl.persistent_blocks[ns]);
    grid = std::min<u64>(grid, 0x7fffffffULL / (u64)(64 * v.wpb));
    if (grid == 0) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: empty grid");
    a.wave_stride = (i64)grid * v.wpb;
    a.slab_shift = pl.slab_shift;
    a.pair_ahead = pl.pair_ahead;
    v.fn(data, a, (unsigned)grid, s, ns);
    return check_launch(v.pair ? "adv_column_pair_kernel" : "adv_column_stream_kernel");
}

// nsteps steps: full passes of the plan's steps per pass, then the rest.
bkma_status plan_steps(const TunedPlan &pl, void *data, u64 nsteps, hipStream_t s) {
    const u64 ns = (u64)pl.ns;
    for (u64 t = 0; t + ns <= nsteps; t += ns) BKMA_TRY(plan_launch(pl, data, s, (int)ns));
    if (nsteps % ns) BKMA_TRY(plan_launch(pl, data, s, (int)(nsteps % ns)));
    return BKMA_OK;
}

std::string stream_name(const StreamVariant &v, int ns, bool persistent) {
    return std::string(v.pair ? "column_pair<" : "column_stream<") + (v.dtype == BKMA_F64 ? "f64" : (v.wide ? "f32,win64" : "f32,win32")) +
           ",vec" + std::to_string(v.vec) + ",smax" + std::to_string(v.smax) + ",p" + std::to_string(v.prefetch) +
           ",wpb" + std::to_string(v.wpb) + ",steps" + std::to_string(ns) + (v.nt ? ",nt" : "") +
           (v.dpp ? ",dpp" : "") + (persistent ? ",persistent" : "") + ">";
}

std::string lds_name(bkma_dtype dt, const LdsArgs &a, int threads) {
    return std::string("line_lds<") + (dt == BKMA_F64 ? "f64" : "f32") +
           (a.nseg > 1 ? ",seg" + std::to_string(a.seg_rows) + "x" + std::to_string(a.nseg) + ",chunk" +
                             std::to_string(a.chunk_segs)
                       : ",whole_line") +
           (a.slabs > 1 ? ",slabs" + std::to_string(a.slabs) : "") + (a.ns > 1 ? ",steps" + std::to_string(a.ns) : "") +
           (threads != 256 ? ",wg" + std::to_string(threads) : "") + ">";
}

// Resident work-groups on the whole device for a variant (persistent grid).
i64 resident_blocks(const StreamVariant &v, int ns, int pair_ahead) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return (i64)v.occ(ns, pair_ahead) * cus;
}

// A candidate plan (not yet owning device memory).
struct Candidate {
    bkma_tuned_config cfg;
};

bkma_status build_plan(const bkma_adv_params &p, bkma_dtype dt, const bkma_tuned_config &cfg, TunedPlan *pl) {
    std::string why;
    pl->params = p;
    pl->dtype = dt;
    pl->cfg = cfg;
    if (cfg.strategy == BKMA_STRATEGY_COLUMN_STREAM || cfg.strategy == BKMA_STRATEGY_COLUMN_PAIR) {
        const int ns = steps_per_pass(cfg);
        const int pair = cfg.strategy == BKMA_STRATEGY_COLUMN_PAIR;
        for (const auto &v : stream_variants(dt)) {
            if (v.pair == pair && v.vec == cfg.vec && v.smax == cfg.smax && v.prefetch == cfg.prefetch &&
                v.wpb == cfg.waves_per_wg && v.nt == (cfg.nontemporal != 0) && v.wide == (cfg.wide_window != 0) &&
                v.dpp == (cfg.dpp_broadcast != 0) && ns <= v.nsmax) {
                if (pair) {
                    int stats[kShiftStats];
                    if (!shift_supported(p, &why))
                        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
                    BKMA_TRY(scan_slab_shifts(p, &pl->slab_shift, stats));
                    if (pair_feasible(p, dt, v, ns, stats, &why) != BKMA_OK)
                        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
                    pl->pair_ahead = std::max(0, 3 - stats[0]);
                } else if (stream_feasible(p, dt, v, ns, &why) != BKMA_OK) {
                    return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
                }
                pl->variant = &v;
                pl->ns = ns;
                pl->cfg.steps_per_pass = ns;
                for (int k = 1; k <= ns; ++k)
                    pl->persistent_blocks[k] = cfg.persistent ? resident_blocks(v, k, pl->pair_ahead) : 0;
                pl->cfg.persistent = pl->persistent_blocks[ns] > 0;
                pl->name = stream_name(v, ns, pl->persistent_blocks[ns] > 0);
                return BKMA_OK;
            }
        }
        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: no compiled variant matches the configuration");
    }
    if (cfg.strategy == BKMA_STRATEGY_PLAIN) {
        // FP64 only: the plain executor's FP32 storage sums with separately
        // rounded multiply and add, the tuned FP32 plans with fused ones
        if (dt != BKMA_F64) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: PLAIN runs FP64 storage only");
        if (steps_per_pass(cfg) != 1) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: PLAIN runs one step per pass");
        pl->cfg.steps_per_pass = 1;
        BKMA_TRY(bkma_create_optim_params(p.n0, p.n1, p.n2, (u64)cfg.seg_rows, 1, 1, 8, 0, &pl->optim));
        if (pl->optim.mem_space != BKMA_MEM_LOCAL)
            return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: PLAIN needs the line in LDS");
        pl->name = "plain_adaptivewg<pref" + std::to_string(cfg.seg_rows) + ">";
        return BKMA_OK;
    }
    if (cfg.strategy == BKMA_STRATEGY_HYBRID) {
        // HybridMem (src/core/experiments/HybridMem.cpp:4-32, its constructor
        // :56-98): a fraction percent_loc of the slabs keeps each line's
        // scratch in LDS, the rest use a global scratch with work-groups sized
        // without the LDS limit; k_local = floor(n0 * percent_loc)
        // (dispatch_kernels, src/core/bkma_tools.hpp:157-165)
        if (dt != BKMA_F64) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: HYBRID runs FP64 storage only");
        if (steps_per_pass(cfg) != 1) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: HYBRID runs one step per pass");
        if (cfg.seg_rows <= 0 || cfg.chunk_segs < 0 || cfg.chunk_segs > 100)
            return fail(BKMA_ERR_INVALID_ARGUMENT,
                        "tuned advection: HYBRID needs seg_rows = pref_wg_size > 0 and chunk_segs = percent in [0, 100]");
        pl->cfg.steps_per_pass = 1;
        pl->k_local = p.n0 * (u64)cfg.chunk_segs / 100;
        const u64 kg = p.n0 - pl->k_local;
        if (pl->k_local > 0) {
            BKMA_TRY(bkma_create_optim_params(pl->k_local, p.n1, p.n2, (u64)cfg.seg_rows, 1, 1, 8, 0, &pl->optim));
            if (pl->optim.mem_space != BKMA_MEM_LOCAL)
                return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: HYBRID needs the LDS slabs' lines in LDS");
        }
        if (kg > 0) {
            // set_ideal_sizes + adjust_sizes_mem_limit(INT_MAX, n1): no LDS limit
            BKMA_TRY(bkma_create_optim_params(kg, p.n1, p.n2, (u64)cfg.seg_rows, 1, 1, 8, (u64)INT32_MAX * 8,
                                              &pl->optim_global));
            pl->optim_global.mem_space = BKMA_MEM_GLOBAL;
            uint64_t e0, e1, e2;
            BKMA_TRY(bkma_scratch_extents(BKMA_IMPL_ADAPTIVE_WG, &pl->optim_global, kg, p.n1, p.n2, 1, &e0, &e1, &e2));
            if (hipMalloc(&pl->gscratch.data, std::max<u64>(e0 * e1 * e2 * 8, 16)) != hipSuccess) {
                pl->gscratch.data = nullptr;
                return fail(BKMA_ERR_OUT_OF_MEMORY, "tuned advection: HYBRID global scratch allocation failed");
            }
            pl->gscratch.e0 = e0;
            pl->gscratch.e1 = e1;
            pl->gscratch.e2 = e2;
        }
        if (pl->k_local > 0 && kg > 0) {
            BKMA_HIP_TRY(hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking));
            BKMA_HIP_TRY(hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming));
            BKMA_HIP_TRY(hipEventCreateWithFlags(&pl->join, hipEventDisableTiming));
        }
        pl->name = "hybrid_adaptivewg<pref" + std::to_string(cfg.seg_rows) + ",local" +
                   std::to_string(cfg.chunk_segs) + "%>";
        return BKMA_OK;
    }
    if (cfg.strategy == BKMA_STRATEGY_LINE_LDS) {
        const int ns = steps_per_pass(cfg);
        if (ns > kLdsMaxSteps)
            return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: LINE_LDS fuses at most 4 steps per pass");
        pl->cfg.steps_per_pass = ns;
        pl->ns = ns;
        const int wpb = cfg.waves_per_wg > 0 ? cfg.waves_per_wg : 4;
        if (wpb != 4 && wpb != 8 && wpb != 16)
            return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: LINE_LDS work-groups have 4, 8 or 16 wavefronts");
        pl->lds_threads = 64 * wpb;
        pl->cfg.waves_per_wg = wpb;
        if (lds_shape(p, dt, cfg.seg_rows, cfg.chunk_segs, ns, &pl->lds, &pl->lds_bytes, &why) != BKMA_OK)
            return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
        if (pl->lds.nseg > 1) {
            const u64 esz = dt == BKMA_F64 ? 8 : 4;
            const u64 bytes =
                p.n0 * (u64)pl->lds.nchunk * (u64)pl->lds.ns * (u64)(pl->lds.hb_max + pl->lds.ha_max) * p.n2 * esz;
            if (hipMalloc(&pl->halo, std::max<u64>(bytes, 16)) != hipSuccess) {
                pl->halo = nullptr;
                return fail(BKMA_ERR_OUT_OF_MEMORY, "tuned advection: halo buffer allocation failed");
            }
        }
        pl->cfg.seg_rows = pl->lds.nseg > 1 ? (int32_t)pl->lds.seg_rows : 0;
        pl->cfg.chunk_segs = (int32_t)(pl->lds.nseg > 1 ? pl->lds.chunk_segs : pl->lds.slabs);
        pl->name = lds_name(dt, pl->lds, pl->lds_threads);
        return BKMA_OK;
    }
    if (cfg.strategy == BKMA_STRATEGY_LINE_SHFL) {
        const int ns = steps_per_pass(cfg);
        int stats[kShiftStats];
        if (!shift_supported(p, &why))
            return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
        const int cells = cfg.vec > 0 ? cfg.vec : 8;
        BKMA_TRY(scan_slab_shifts(p, &pl->slab_shift, stats));
        if (shfl_feasible(p, cells, ns, stats, &pl->shfl_waves, &why, &pl->row_wmin, &pl->row_tw) != BKMA_OK)
            return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: configuration infeasible (" + why + ")");
        pl->ns = ns;
        pl->cfg.steps_per_pass = ns;
        pl->cfg.vec = cells;
        pl->cfg.waves_per_wg = pl->shfl_waves;
        pl->name = std::string("line_shfl<") + (dt == BKMA_F64 ? "f64" : "f32") +
                   (cells == 1 ? std::string(",row16")
                               : ",cells" + std::to_string(cells) + ",wg" + std::to_string(64 * pl->shfl_waves)) +
                   ",steps" + std::to_string(ns) + ">";
        return BKMA_OK;
    }
    return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: unknown strategy");
}

// Every feasible configuration: compiled column-stream variants (both grid
// modes) and LINE_LDS with whole lines or segments sized for 32/64/160 KB.
std::vector<bkma_tuned_config> candidates(const bkma_adv_params &p, bkma_dtype dt, std::string *why) {
    std::vector<bkma_tuned_config> out;
    // slab shift statistics for the column-pair variants (one device scan)
    int stats[kShiftStats] = {0, 0, 2, 0, 0};
    bool have_stats = false;
    for (const auto &v : stream_variants(dt))
        if (v.pair && !have_stats && shift_supported(p, nullptr)) {
            have_stats = scan_slab_shifts(p, nullptr, stats) == BKMA_OK;
            if (!have_stats) stats[2] = 2;
        }
    auto feasible = [&](const StreamVariant &v, int ns) {
        return v.pair ? pair_feasible(p, dt, v, ns, stats, why) == BKMA_OK
                      : stream_feasible(p, dt, v, ns, why) == BKMA_OK;
    };
    int smin_needed = 1 << 30;
    for (const auto &v : stream_variants(dt))
        if (feasible(v, 1)) smin_needed = std::min(smin_needed, v.smax);
    for (const auto &v : stream_variants(dt)) {
        if (v.smax > smin_needed + 2) continue;
        for (int ns = 1; ns <= v.nsmax; ++ns) {
            if (!feasible(v, ns)) continue;
            for (int pers = 0; pers < 2; ++pers) {
                bkma_tuned_config c;
                std::memset(&c, 0, sizeof(c));
                c.strategy = v.pair ? BKMA_STRATEGY_COLUMN_PAIR : BKMA_STRATEGY_COLUMN_STREAM;
                c.vec = v.vec;
                c.smax = v.smax;
                c.prefetch = v.prefetch;
                c.waves_per_wg = v.wpb;
                c.nontemporal = v.nt;
                c.wide_window = v.wide;
                c.dpp_broadcast = v.dpp;
                c.steps_per_pass = ns;
                c.persistent = pers;
                out.push_back(c);
            }
        }
    }
    // contiguous lines in registers, neighbours through DPP
    for (int cells : {8, 4, 1}) {
        int nw = 0;
        if (shfl_feasible(p, cells, 1, stats, &nw, why) != BKMA_OK) continue;
        for (int ns : {1, 4, 16, 64, 256}) {
            bkma_tuned_config c;
            std::memset(&c, 0, sizeof(c));
            c.strategy = BKMA_STRATEGY_LINE_SHFL;
            c.vec = cells;
            c.waves_per_wg = nw;
            c.steps_per_pass = ns;
            out.push_back(c);
        }
    }
    const u64 esz = dt == BKMA_F64 ? 8 : 4;
    LdsArgs a;
    u64 bytes, lds_max = 0;
    if (lds_limit_bytes(&lds_max) != BKMA_OK) lds_max = 0;
    int dlo = 0, dhi = 0;
    if (params_finite(p)) shift_range(p, &dlo, &dhi);
    const i64 halo_rows = std::max(0, dhi + 2) + std::max(0, 3 - dlo);
    for (int ns = 1; ns <= kLdsMaxSteps; ++ns) {
        // whole lines, and segments sized for 32/64/160 KB of LDS: staged rows
        // in the storage type plus, for fused steps, two FP64 work buffers
        const u64 row_bytes = std::max<u64>(p.n2 * (esz + (ns > 1 ? 16 : 0)), 1);
        std::vector<i64> segs = {0};
        for (u64 budget : {32768ull, 65536ull, 163840ull}) {
            const i64 rows = (i64)(budget / row_bytes) - ns * halo_rows - 1;
            if (rows >= 32 && (u64)rows < p.n1) segs.push_back(rows & ~(i64)3);
        }
        for (i64 sr : segs) {
            std::vector<i64> css = {1, 4, 16, 64};
            if (sr == 0 && p.n1 * p.n2 < 4096) css.push_back((i64)(4096 / (p.n1 * p.n2)));  // slab groups
            for (i64 cs : css) {
                if (lds_shape(p, dt, sr, cs, ns, &a, &bytes, why) != BKMA_OK) continue;
                bkma_tuned_config c;
                std::memset(&c, 0, sizeof(c));
                c.strategy = BKMA_STRATEGY_LINE_LDS;
                c.steps_per_pass = ns;
                c.seg_rows = (int32_t)(a.nseg > 1 ? a.seg_rows : 0);
                c.chunk_segs = (int32_t)(a.nseg > 1 ? a.chunk_segs : a.slabs);
                // a work-group that takes over a quarter of the LDS leaves
                // few waves per CU at 4 per work-group: 8 and 16 too
                for (int wpb : {4, 8, 16}) {
                    if (wpb > 4 && bytes * 4 <= lds_max) break;
                    if (wpb == 16 && ns > 1) break;  // the fused-step kernel spills at 1024 threads
                    c.waves_per_wg = wpb;
                    bool dup = false;
                    for (auto &o : out)
                        dup |= o.strategy == c.strategy && o.seg_rows == c.seg_rows && o.chunk_segs == c.chunk_segs &&
                               o.steps_per_pass == c.steps_per_pass && o.waves_per_wg == c.waves_per_wg;
                    if (!dup) out.push_back(c);
                }
            }
        }
    }
    // the reference's AdaptiveWg executor and dispatch policy (FP64 storage,
    // line in LDS): plans never time slower than it
    if (dt == BKMA_F64) {
        for (int pref : {128, 256, 1024}) {
            bkma_optim_params o;
            if (bkma_create_optim_params(p.n0, p.n1, p.n2, (u64)pref, 1, 1, 8, 0, &o) != BKMA_OK ||
                o.mem_space != BKMA_MEM_LOCAL)
                continue;
            bkma_tuned_config c;
            std::memset(&c, 0, sizeof(c));
            c.strategy = BKMA_STRATEGY_PLAIN;
            c.steps_per_pass = 1;
            c.seg_rows = pref;
            out.push_back(c);
        }
        // local/global splits of the slabs (HybridMem), at pref_wg_size 256
        for (int pct : {25, 50, 75}) {
            const u64 kl = p.n0 * (u64)pct / 100;
            bkma_optim_params o;
            if (kl == 0 || kl == p.n0 || bkma_create_optim_params(kl, p.n1, p.n2, 256, 1, 1, 8, 0, &o) != BKMA_OK ||
                o.mem_space != BKMA_MEM_LOCAL)
                continue;
            bkma_tuned_config c;
            std::memset(&c, 0, sizeof(c));
            c.strategy = BKMA_STRATEGY_HYBRID;
            c.steps_per_pass = 1;
            c.seg_rows = 256;
            c.chunk_segs = pct;
            out.push_back(c);
        }
    }
    return out;
}

// Heuristic (measured on the strided 2048x1024x512 workload): column pair
// when every slab has at most two shifts (FP32 with the DPP broadcast at 5,
// FP64 at 4 steps per pass), else the column stream with the smallest window (FP32 at 2, FP64 at
// 3 steps per pass; more steps push the register window past the occupancy
// that hides HBM latency), widest vector, deepest prefetch; contiguous
// lines that fit LINE_SHFL with 256 steps per pass, 8 cells per lane from
// 4096 lines on (measured on [2^21][1024][1]), else 4 (on [256][1024][1]);
// else LINE_LDS whole line, else the longest segment.
int heuristic_score(const bkma_tuned_config &c, bkma_dtype dt, const bkma_adv_params &p) {
    if (c.strategy == BKMA_STRATEGY_PLAIN)  // short lines: one work-group per slab would idle
        return (p.n1 < 64 ? 900000000 : 10000000) - std::abs(c.seg_rows - 256);
    if (c.strategy == BKMA_STRATEGY_HYBRID) return 5000000 + c.chunk_segs;
    if (c.strategy == BKMA_STRATEGY_LINE_SHFL) {
        // 256 steps per pass; 8 cells per lane when the lines fill the GPU
        // several times over, else 4 (twice the wavefronts)
        const bool many = p.n0 >= 4096;
        return 600000000 + ((c.vec == 8) == many ? 1000 : 0) - std::abs(steps_per_pass(c) - 256);
    }
    if ((c.strategy == BKMA_STRATEGY_COLUMN_PAIR || c.strategy == BKMA_STRATEGY_COLUMN_STREAM) &&
        p.n2 < 64 * (u64)c.vec)  // a wavefront's lanes would mostly idle
        return 1000 + c.vec;
    if (c.strategy == BKMA_STRATEGY_COLUMN_PAIR) {
        // FP32: DPP broadcast, FP64 stage-0 window, VEC 2, P 14, 5 steps;
        // FP64: VEC 2, P 14, 4 steps
        const int want = dt == BKMA_F64 ? 4 : (c.dpp_broadcast ? 5 : 3);
        const int ns = steps_per_pass(c);
        return 1100000000 + c.dpp_broadcast * 10000000 - c.smax * 1000000 - std::abs(ns - want) * 100000 +
               500 * c.wide_window + c.prefetch * 10 + c.nontemporal - c.persistent +
               (c.dpp_broadcast ? -c.vec : c.vec) * 1000;
    }
    if (c.strategy == BKMA_STRATEGY_COLUMN_STREAM) {
        const int want = dt == BKMA_F64 ? 3 : 2;
        const int ns = steps_per_pass(c);
        return 1000000000 - c.smax * 1000000 - std::abs(ns - want) * 100000 + c.vec * 1000 + c.prefetch * 10 +
               c.nontemporal - c.persistent - 100 * c.wide_window;
    }
    // LINE_LDS: whole lines; else the longest segment, with 16-wave
    // work-groups where they are offered (the segment fills the LDS)
    return c.seg_rows == 0 ? 500000000 - std::abs(steps_per_pass(c) - 2) * 1000 - c.waves_per_wg
                           : 100000000 + std::min(c.seg_rows, 1 << 24) / 1024 - std::abs(c.chunk_segs - 16) * 10 +
                                 (c.waves_per_wg == 16 ? 5 : 0);
}

bkma_status time_plan(const TunedPlan &pl, void *buf, int reps, hipStream_t s, float *ms) {
    hipEvent_t e0, e1;
    BKMA_HIP_TRY(hipEventCreate(&e0));
    BKMA_HIP_TRY(hipEventCreate(&e1));
    bkma_status st = plan_launch(pl, buf, s, pl.ns);  // warm-up
    if (st == BKMA_OK) {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && st == BKMA_OK; ++r) st = plan_launch(pl, buf, s, pl.ns);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(ms, e0, e1);
        *ms /= (float)(reps * pl.ns);  // per step
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

}  // namespace

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

    std::string why;
    std::vector<bkma_tuned_config> cands = candidates(*params, dtype, &why);
    if (cands.empty())
        return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: no strategy covers this shape (" + why + ")");
    bkma_tuned_config best = *std::max_element(cands.begin(), cands.end(), [&](const auto &a, const auto &b) {
        return heuristic_score(a, dtype, *params) < heuristic_score(b, dtype, *params);
    });

    if (autotune && cands.size() > 1) {
        // Time the candidates on a scratch array of whole slabs, large enough
        // to fill the GPU several times over (contents do not affect timing).
        // The whole problem (up to 2 Gi cells) is used: occupancy and tail
        // effects depend on the number of work units.
        bkma_adv_params sub = *params;
        const u64 elems_per_slab = params->n1 * params->n2;
        const u64 target = 2048ull << 20;
        sub.n0 = std::max<u64>(1, std::min<u64>(params->n0, target / std::max<u64>(elems_per_slab, 1)));
        const size_t bytes = (size_t)(sub.n0 * elems_per_slab) * (dtype == BKMA_F64 ? 8 : 4);
        void *buf = nullptr;
        if (hipMalloc(&buf, bytes) != hipSuccess) {  // no room for the scratch: keep the heuristic choice
            (void)hipGetLastError();
            buf = nullptr;
        }
        if (buf) (void)hipMemsetAsync(buf, 0, bytes, as_stream(stream));
        float best_ms = 1e30f;
        for (const auto &c : cands) {
            if (!buf) break;
            TunedPlan pl;
            if (build_plan(sub, dtype, c, &pl) != BKMA_OK) continue;
            float ms = 0;
            if (time_plan(pl, buf, 3, as_stream(stream), &ms) == BKMA_OK && ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        (void)hipFree(buf);
        (void)hipGetLastError();
    }
    auto *pl = new (std::nothrow) bkma_adv_plan();
    if (!pl) return fail(BKMA_ERR_OUT_OF_MEMORY, "plan allocation");
    bkma_status st = build_plan(*params, dtype, best, pl);
    if (st != BKMA_OK) {
        delete pl;
        return st;
    }
    *out = pl;
    return BKMA_OK;
}

extern "C" bkma_status bkma_adv_plan_create_with(const bkma_adv_params *params, bkma_dtype dtype,
                                                 const bkma_tuned_config *cfg, bkma_adv_plan **out) {
    if (!params || !cfg || !out) return fail(BKMA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (dtype != BKMA_F32 && dtype != BKMA_F64) return fail(BKMA_ERR_UNSUPPORTED, "tuned advection: F32 or F64");
    if (params->n0 == 0 || params->n1 == 0 || params->n2 == 0)
        return fail(BKMA_ERR_INVALID_ARGUMENT, "array extents must be positive");
    BKMA_TRY(require_device());
    auto *pl = new (std::nothrow) bkma_adv_plan();
    if (!pl) return fail(BKMA_ERR_OUT_OF_MEMORY, "plan allocation");
    bkma_status st = build_plan(*params, dtype, *cfg, pl);
    if (st != BKMA_OK) {
        delete pl;
        return st;
    }
    *out = pl;
    return BKMA_OK;
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
    return plan_steps(*plan, data, nsteps, as_stream(stream));
}

extern "C" void bkma_adv_plan_destroy(bkma_adv_plan *plan) { delete plan; }

extern "C" int32_t bkma_adv_plan_variants(bkma_dtype dtype, bkma_tuned_config *out, int32_t max_out) {
    int32_t n = 0;
    for (const auto &v : stream_variants(dtype)) {
        for (int ns = 1; ns <= v.nsmax; ++ns) {
            if (out && n < max_out) {
                bkma_tuned_config c;
                std::memset(&c, 0, sizeof(c));
                c.strategy = v.pair ? BKMA_STRATEGY_COLUMN_PAIR : BKMA_STRATEGY_COLUMN_STREAM;
                c.vec = v.vec;
                c.smax = v.smax;
                c.prefetch = v.prefetch;
                c.waves_per_wg = v.wpb;
                c.nontemporal = v.nt;
                c.wide_window = v.wide;
                c.dpp_broadcast = v.dpp;
                c.steps_per_pass = ns;
                out[n] = c;
            }
            ++n;
        }
    }
    return n;
}
