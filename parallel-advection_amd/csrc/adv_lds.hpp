// Device side of the LDS-staged advection kernels (see bkma_tuned.hip).
//
// A work unit is one segment [r0, r1) of the rows of one i0 slab (all n2
// columns).  The segment and the halo rows its stencils reach outside it are
// staged in LDS, then every output of the segment is computed from LDS and
// written once to HBM.
//  - Whole-line mode (one segment per slab): the halo is the line's own wrap
//    around, read with modular indexing; nothing else is needed.  This mode
//    can fuse ns steps per pass (line_steps_unit): the staged line is
//    advanced ns times inside LDS, the intermediate steps in FP64 work
//    buffers, and only the last step is written to HBM.
//  - Segmented mode (lines longer than the LDS budget): other work-groups
//    overwrite the rows around a segment concurrently, so the halo rows are
//    first copied to a scratch buffer by halo_snapshot_kernel; the halo of
//    slab i0 only spans the rows its own velocity reaches (the stencil of
//    output i1 reads rows i1-d-2 .. i1-d+3 with d in [floor(s), floor(s)+1],
//    s = v(i0)*dt/dx).
#pragma once

#include "common.hpp"

namespace bkma {

struct LdsArgs {
    AdvDev p;
    i64 seg_rows;    // rows per segment (== n1 in whole-line mode)
    i64 nseg;        // segments per slab
    i64 hb_max;      // halo rows before a segment (allocation bound)
    i64 ha_max;      // halo rows after a segment (allocation bound)
    i64 chunk_segs;  // segmented mode: consecutive segments one work-group walks in order
    i64 nchunk;      // chunks per slab
    i64 ns;          // whole-line mode: steps per pass
    i64 slabs;       // whole-line mode: consecutive slabs one unit stages together (short lines)
};

// Halo extents of slab i0: rows [r0 - hb, r0) and [r1, r1 + ha).
__device__ __forceinline__ void slab_halo(const AdvDev &p, i64 i0, i64 hb_max, i64 ha_max, int &hb, int &ha) {
    const double v = adv_coord((int)i0, p.minRealVx, p.dvx);
    const double s = v * p.dt * p.inv_dx;
    const double dlo = floor(s - 1e-6), dhi = floor(s + 1e-6) + 1.0;
    hb = (int)fmin((double)hb_max, fmax(0.0, dhi + 2.0));
    ha = (int)fmin((double)ha_max, fmax(0.0, 3.0 - dlo));
}

// Copies, for every (slab, segment), the old halo rows into
// halo[(i0 * nseg + seg) * ns * (hb_max + ha_max) * n2 ...]: first the ns*hb
// rows before the segment, then the ns*ha rows after it (periodic in i1).
template <class T>
__device__ __forceinline__ void halo_snapshot_unit(const T *__restrict__ data, T *__restrict__ halo,
                                                   const LdsArgs &a, i64 unit) {
    const AdvDev &p = a.p;
    const i64 i0 = unit / a.nseg, seg = unit - i0 * a.nseg;
    const i64 n1 = p.n1, n2 = p.n2;
    const i64 r0 = seg * a.seg_rows;
    const i64 r1 = min(n1, r0 + a.seg_rows);
    int hb, ha;
    slab_halo(p, i0, a.hb_max, a.ha_max, hb, ha);
    hb *= (int)a.ns;  // ns fused steps reach ns times as far
    ha *= (int)a.ns;
    const T *slab = data + i0 * n1 * n2;
    T *dst = halo + unit * a.ns * (a.hb_max + a.ha_max) * n2;
    const i64 total = (i64)(hb + ha) * n2;
    for (i64 e = threadIdx.x; e < total; e += blockDim.x) {
        const i64 hr = e / n2, c = e - hr * n2;
        i64 row = hr < hb ? r0 - hb + hr : r1 + (hr - hb);
        row %= n1;
        if (row < 0) row += n1;
        dst[e] = slab[row * n2 + c];
    }
}

// Grid-stride over the (slab, segment) units: a grid is capped at
// kLdsMaxGrid work-groups (the launch limit is 2^32 work-items).
constexpr i64 kLdsMaxGrid = (i64)1 << 22;
template <class T>
__global__ void halo_snapshot_kernel(const T *__restrict__ data, T *__restrict__ halo, LdsArgs a) {
    for (i64 unit = blockIdx.x; unit < a.p.n0 * a.nseg; unit += gridDim.x) halo_snapshot_unit(data, halo, a, unit);
}

// Outputs of rows [r0, r0+nr) from the staged rows in LDS.  Thread (row
// offset, chunk of V columns); the foot location and coefficients are
// computed once per (row, chunk); chunks beyond the work-group loop.
template <class T, bool SEG, int V>
__device__ __forceinline__ void lds_compute(T *__restrict__ slab, const AdvDev &p, i64 i0, i64 r0, i64 nr, int hb,
                                            i64 lds_rows, const T *__restrict__ lds) {
    using VT = T __attribute__((ext_vector_type(V)));
    const i64 n1 = p.n1, n2 = p.n2;
    const int nthr = blockDim.x;
    const i64 nch = n2 / V;
    const int cw = (int)min<i64>(nch, nthr);
    const int rows_per_pass = nthr / cw;
    const int t_row = threadIdx.x / cw, t_col = threadIdx.x - t_row * cw;
    const int half = (int)(n1 / 2);
    double coef[6];
    unsigned long long cached = ~0ull;
    for (i64 rr = t_row; t_row < rows_per_pass && rr < nr; rr += rows_per_pass) {
        const i64 i1 = r0 + rr;
        double d_prev1;
        const int leftNode = adv_foot(p, (int)i0, (int)i1, d_prev1);
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, d_prev1);
        if (bits != cached) {
            adv_lag_basis(d_prev1, coef);
            cached = bits;
        }
        const int ipos1 = leftNode - 2;
        i64 lrow;  // LDS row of the first stencil point
        bool ok = true;
        if constexpr (SEG) {
            int delta = ipos1 - (int)i1;
            if (delta < -half) delta += (int)n1;
            if (delta > half) delta -= (int)n1;
            lrow = hb + rr + delta;
            ok = lrow >= 0 && lrow + 5 < lds_rows;
        } else {
            lrow = stencil_fast_ok(n1, ipos1) ? stencil_base(n1, ipos1) : stencil_index_slow(n1, ipos1, 0);
        }
        for (i64 ch = t_col; ch < nch; ch += cw) {
            double acc[V];
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] = 0.;
            i64 r = lrow;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                T f[V];
                if constexpr (V == 1) {
                    f[0] = lds[r * n2 + ch];
                } else {
                    const VT fv = *reinterpret_cast<const VT *>(lds + r * n2 + ch * V);
#pragma unroll
                    for (int v = 0; v < V; ++v) f[v] = fv[v];
                }
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const double fd = ok ? (double)f[v] : (double)NAN;
                    if constexpr (sizeof(T) == 8) acc[v] = acc[v] + coef[k] * fd;
                    else acc[v] = fma(coef[k], fd, acc[v]);
                }
                r += 1;
                if constexpr (!SEG) {
                    if (r >= n1) r -= n1;
                }
            }
            if constexpr (V == 1) {
                slab[i1 * n2 + ch] = (T)acc[0];
            } else {
                VT o;
#pragma unroll
                for (int v = 0; v < V; ++v) o[v] = (T)acc[v];
                *reinterpret_cast<VT *>(slab + i1 * n2 + ch * V) = o;
            }
        }
    }
}

// One LDS-staged unit: rows [r0, r1) of slab i0 (all n2 columns).  Every
// thread of the work-group calls it; it contains one barrier and no early
// return, so it can run inside loops of a persistent kernel.
template <class T, bool SEG>
__device__ __forceinline__ void line_lds_unit(T *__restrict__ slab, const T *__restrict__ hsrc, const AdvDev &p, i64 i0,
                                              i64 r0, i64 r1, i64 hb_max, i64 ha_max, T *__restrict__ lds) {
    const i64 n2 = p.n2;
    const i64 nr = r1 - r0;
    const int nthr = blockDim.x;
    int hb = 0, ha = 0;
    if constexpr (SEG) slab_halo(p, i0, hb_max, ha_max, hb, ha);
    const i64 lds_rows = hb + nr + ha;

    // Stage: the segment (contiguous in HBM), then the halo rows.
    {
        const i64 cnt = nr * n2;
        const T *src = slab + r0 * n2;
        T *dst = lds + (i64)hb * n2;
        constexpr int V = 16 / sizeof(T);
        const bool vec_ok = ((reinterpret_cast<uintptr_t>(src) | (uintptr_t)(hb * n2 * sizeof(T))) % 16) == 0;
        if (vec_ok) {
            using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
            const i64 nv = cnt / V;
            for (i64 e = threadIdx.x; e < nv; e += nthr)
                reinterpret_cast<VT *>(dst)[e] = reinterpret_cast<const VT *>(src)[e];
            for (i64 e = nv * V + threadIdx.x; e < cnt; e += nthr) dst[e] = src[e];
        } else {
            for (i64 e = threadIdx.x; e < cnt; e += nthr) dst[e] = src[e];
        }
        if constexpr (SEG) {
            for (i64 e = threadIdx.x; e < (i64)hb * n2; e += nthr) lds[e] = hsrc[e];
            for (i64 e = threadIdx.x; e < (i64)ha * n2; e += nthr) lds[(hb + nr) * n2 + e] = hsrc[(i64)hb * n2 + e];
        }
    }
    __syncthreads();

    // Compute: 16-byte column chunks when n2 allows (one foot location per
    // chunk, vector LDS reads and global stores), else single columns.
    if (n2 % (16 / (i64)sizeof(T)) == 0)
        lds_compute<T, SEG, 16 / (int)sizeof(T)>(slab, p, i0, r0, nr, hb, lds_rows, lds);
    else
        lds_compute<T, SEG, 1>(slab, p, i0, r0, nr, hb, lds_rows, lds);
}

// One step of whole staged lines: dst = step(src) for the `ns` consecutive
// slabs i0 .. i0+ns-1, both holding their n1 rows of n2 columns each (src in
// LDS, dst in LDS or the slabs in HBM).  Arithmetic of the storage type T:
// FP64 storage sums with separately rounded multiply and add, FP32 storage
// with fused ones, in FP64 either way.  Thread (row of some slab, group of VC
// columns); the foot is computed once per (row, group).
template <class T, int VC, class S, class D>
__device__ __forceinline__ void line_step(const S *__restrict__ src, D *__restrict__ dst, const AdvDev &p, i64 i0,
                                          i64 nslab) {
    const i64 n1 = p.n1, n2 = p.n2;
    const int nthr = blockDim.x;
    const i64 nch = n2 / VC;
    const int cw = (int)min<i64>(nch, nthr);
    const int rows_per_pass = nthr / cw;
    const int t_row = threadIdx.x / cw, t_col = threadIdx.x - t_row * cw;
    if (t_row >= rows_per_pass) return;
    double coef[6];
    unsigned long long cached = ~0ull;
    for (i64 rr = t_row; rr < nslab * n1; rr += rows_per_pass) {
        const i64 g = rr / n1, i1 = rr - g * n1;
        const S *gsrc = src + g * n1 * n2;
        D *gdst = dst + g * n1 * n2;
        double d_prev1;
        const int leftNode = adv_foot(p, (int)(i0 + g), (int)i1, d_prev1);
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, d_prev1);
        if (bits != cached) {
            adv_lag_basis(d_prev1, coef);
            cached = bits;
        }
        const int ipos1 = leftNode - 2;
        const i64 base = stencil_fast_ok(n1, ipos1) ? stencil_base(n1, ipos1) : stencil_index_slow(n1, ipos1, 0);
        for (i64 ch = t_col; ch < nch; ch += cw) {
            double acc[VC];
#pragma unroll
            for (int v = 0; v < VC; ++v) acc[v] = 0.;
            i64 r = base;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
#pragma unroll
                for (int v = 0; v < VC; ++v) {
                    const double f = (double)gsrc[r * n2 + ch * VC + v];
                    if constexpr (sizeof(T) == 8) acc[v] = acc[v] + coef[k] * f;
                    else acc[v] = fma(coef[k], f, acc[v]);
                }
                r += 1;
                if (r >= n1) r -= n1;
            }
#pragma unroll
            for (int v = 0; v < VC; ++v) gdst[i1 * n2 + ch * VC + v] = (D)acc[v];
        }
    }
}

// The feet of a thread's rows, kept across the fused steps of a pass (the
// foot of row i1 of slab i0 is the same at every step): for each of its at
// most RMAX rows, the first stencil row and the six coefficients.  line_step
// computes the same values at every step; for contiguous lines (n2 = 1) that
// FP64 work dominates the step, one foot per cell.
template <int RMAX> struct LineFeet {
    i64 base[RMAX];
    double coef[RMAX][6];
};

// Thread layout of line_step for VC columns per thread.
struct LineLayout {
    int cw, rows_per_pass, t_row, t_col;
};
template <int VC> __device__ __forceinline__ LineLayout line_layout(const AdvDev &p) {
    LineLayout l;
    const i64 nch = p.n2 / VC;
    l.cw = (int)min<i64>(nch, blockDim.x);
    l.rows_per_pass = blockDim.x / l.cw;
    l.t_row = threadIdx.x / l.cw;
    l.t_col = threadIdx.x - l.t_row * l.cw;
    return l;
}

template <int VC, int RMAX>
__device__ __forceinline__ void line_feet(const AdvDev &p, i64 i0, i64 nslab, LineFeet<RMAX> &ft) {
    const LineLayout l = line_layout<VC>(p);
    const i64 n1 = p.n1;
#pragma unroll
    for (int j = 0; j < RMAX; ++j) {
        const i64 rr = l.t_row + (i64)j * l.rows_per_pass;
        ft.base[j] = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) ft.coef[j][k] = 0.;
        if (l.t_row < l.rows_per_pass && rr < nslab * n1) {
            const i64 g = rr / n1, i1 = rr - g * n1;
            double d_prev1;
            const int leftNode = adv_foot(p, (int)(i0 + g), (int)i1, d_prev1);
            adv_lag_basis(d_prev1, ft.coef[j]);
            const int ipos1 = leftNode - 2;
            ft.base[j] = stencil_fast_ok(n1, ipos1) ? stencil_base(n1, ipos1) : stencil_index_slow(n1, ipos1, 0);
        }
    }
}

// line_step with the feet of LineFeet (every row of the unit is one of the
// thread's at most RMAX rows).
template <class T, int VC, int RMAX, class S, class D>
__device__ __forceinline__ void line_step_feet(const S *__restrict__ src, D *__restrict__ dst, const AdvDev &p,
                                               i64 nslab, const LineFeet<RMAX> &ft) {
    const LineLayout l = line_layout<VC>(p);
    const i64 n1 = p.n1, n2 = p.n2, nch = n2 / VC;
    if (l.t_row >= l.rows_per_pass) return;
#pragma unroll
    for (int j = 0; j < RMAX; ++j) {
        const i64 rr = l.t_row + (i64)j * l.rows_per_pass;
        if (rr >= nslab * n1) break;
        const i64 g = rr / n1, i1 = rr - g * n1;
        const S *gsrc = src + g * n1 * n2;
        D *gdst = dst + g * n1 * n2;
        for (i64 ch = l.t_col; ch < nch; ch += l.cw) {
            double acc[VC];
#pragma unroll
            for (int v = 0; v < VC; ++v) acc[v] = 0.;
            i64 r = ft.base[j];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
#pragma unroll
                for (int v = 0; v < VC; ++v) {
                    const double f = (double)gsrc[r * n2 + ch * VC + v];
                    if constexpr (sizeof(T) == 8) acc[v] = acc[v] + ft.coef[j][k] * f;
                    else acc[v] = fma(ft.coef[j][k], f, acc[v]);
                }
                r += 1;
                if (r >= n1) r -= n1;
            }
#pragma unroll
            for (int v = 0; v < VC; ++v) gdst[i1 * n2 + ch * VC + v] = (D)acc[v];
        }
    }
}

// Rows per thread whose feet a fused pass keeps in registers (about 40
// VGPRs per row): none for 16-wave work-groups, whose waves have 128 VGPRs.
template <int NT> constexpr int line_feet_rows() { return NT <= 256 ? 4 : (NT <= 512 ? 2 : 0); }

// Whole lines of the nslab consecutive slabs i0 .. advanced by ns steps in
// one pass: staged once into `lds` (storage type), steps 1 .. ns-1 into the
// FP64 buffers w0, w1 in turn, step ns from LDS to the slabs in HBM.  When
// every thread has at most RMAX rows, the feet are computed once for the ns
// steps.  Every thread calls it; it has no early return.
template <class T, int RMAX>
__device__ __forceinline__ void line_steps_unit(T *__restrict__ slab, const AdvDev &p, i64 i0, i64 nslab, int ns,
                                                T *__restrict__ lds, double *__restrict__ w0,
                                                double *__restrict__ w1) {
    const i64 cnt = nslab * p.n1 * p.n2;
    for (i64 e = threadIdx.x; e < cnt; e += blockDim.x) lds[e] = slab[e];
    const bool even = p.n2 % 2 == 0;
    constexpr int RM = RMAX > 0 ? RMAX : 1;
    LineFeet<RM> ft;
    bool keep = false;  // uniform
    if constexpr (RMAX > 0) {
        const LineLayout lay = even ? line_layout<2>(p) : line_layout<1>(p);
        keep = ns > 1 && nslab * p.n1 <= (i64)RMAX * lay.rows_per_pass;
        if (keep) {
            if (even)
                line_feet<2>(p, i0, nslab, ft);
            else
                line_feet<1>(p, i0, nslab, ft);
        }
    }
    __syncthreads();
    auto step = [&](const auto *src, auto *dst) __attribute__((always_inline)) {
        if (RMAX > 0 && keep) {
            if (even)
                line_step_feet<T, 2>(src, dst, p, nslab, ft);
            else
                line_step_feet<T, 1>(src, dst, p, nslab, ft);
        } else if (even) {
            line_step<T, 2>(src, dst, p, i0, nslab);
        } else {
            line_step<T, 1>(src, dst, p, i0, nslab);
        }
        __syncthreads();
    };
    if (ns == 1) {
        step(lds, slab);
        return;
    }
    step(lds, w0);
    for (int t = 2; t < ns; ++t) {
        if (t % 2 == 0)
            step((const double *)w0, w1);
        else
            step((const double *)w1, w0);
    }
    step((const double *)(ns % 2 == 0 ? w0 : w1), slab);
}

template <class T, bool SEG, int NT>
__global__ __launch_bounds__(NT) void adv_line_lds_kernel(T *__restrict__ data, const T *__restrict__ halo,
                                                          LdsArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (!SEG && (a.ns > 1 || a.slabs > 1)) {
        // whole lines of `slabs` slabs, ns steps: [slabs][n1][n2] T, then two
        // such FP64 buffers
        const i64 per = a.p.n1 * a.p.n2, cnt = a.slabs * per;
        T *st = reinterpret_cast<T *>(lds_raw);
        double *w0 = reinterpret_cast<double *>(lds_raw + ((cnt * (i64)sizeof(T) + 15) & ~(i64)15));
        const i64 units = (a.p.n0 + a.slabs - 1) / a.slabs;
        for (i64 unit = blockIdx.x; unit < units; unit += gridDim.x) {
            const i64 i0 = unit * a.slabs;
            const i64 nslab = min(a.slabs, a.p.n0 - i0);
            line_steps_unit<T, line_feet_rows<NT>()>(data + i0 * per, a.p, i0, nslab, (int)a.ns, st, w0, w0 + cnt);
            __syncthreads();  // the next unit restages the LDS
        }
        return;
    }
    for (i64 unit = blockIdx.x; unit < a.p.n0 * a.nseg; unit += gridDim.x) {
        const i64 i0 = unit / a.nseg, seg = unit - i0 * a.nseg;
        const i64 r0 = seg * a.seg_rows;
        const i64 r1 = min(a.p.n1, r0 + a.seg_rows);
        {
            line_lds_unit<T, SEG>(data + i0 * a.p.n1 * a.p.n2,
                                  SEG ? halo + unit * (a.hb_max + a.ha_max) * a.p.n2 : nullptr, a.p, i0, r0, r1,
                                  a.hb_max, a.ha_max, reinterpret_cast<T *>(lds_raw));
        }
        __syncthreads();  // the next unit restages the LDS
    }
}

// Rows [c0, c1) of slab i0 ("chunk"), walked by one work-group in segments
// of seg_rows in increasing order.  Only the chunk's own rows are read from
// HBM besides the halo snapshot of the chunk (hsnap: hb rows before c0, then
// ha rows after c1): the halo rows before a later segment are old rows of the
// previous segment, still in LDS, moved to the front (seg_rows >= hb_max);
// the rows after a segment are read from HBM while they are still unwritten.
// Every thread calls it; the loop is uniform over the work-group.
template <class T>
__device__ __forceinline__ void line_chunk_unit(T *__restrict__ slab, const T *__restrict__ hsnap, const AdvDev &p,
                                                i64 i0, i64 c0, i64 c1, i64 seg_rows, i64 hb_max, i64 ha_max,
                                                T *__restrict__ lds) {
    const i64 n2 = p.n2;
    const int nthr = blockDim.x;
    int hb, ha;
    slab_halo(p, i0, hb_max, ha_max, hb, ha);
    for (i64 r0 = c0; r0 < c1; r0 += seg_rows) {
        const i64 r1 = min(c1, r0 + seg_rows);
        const i64 nr = r1 - r0;
        const i64 lds_rows = hb + nr + ha;
        if (r0 == c0)
            for (i64 e = threadIdx.x; e < (i64)hb * n2; e += nthr) lds[e] = hsnap[e];
        {
            const i64 cnt = nr * n2;
            const T *src = slab + r0 * n2;
            T *dst = lds + (i64)hb * n2;
            constexpr int V = 16 / sizeof(T);
            const bool vec_ok = ((reinterpret_cast<uintptr_t>(src) | (uintptr_t)(hb * n2 * sizeof(T))) % 16) == 0;
            if (vec_ok) {
                using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
                const i64 nv = cnt / V;
                for (i64 e = threadIdx.x; e < nv; e += nthr)
                    reinterpret_cast<VT *>(dst)[e] = reinterpret_cast<const VT *>(src)[e];
                for (i64 e = nv * V + threadIdx.x; e < cnt; e += nthr) dst[e] = src[e];
            } else {
                for (i64 e = threadIdx.x; e < cnt; e += nthr) dst[e] = src[e];
            }
            T *adst = lds + (hb + nr) * n2;
            for (i64 e = threadIdx.x; e < (i64)ha * n2; e += nthr) {
                const i64 g = r1 + e / n2;  // global row
                adst[e] = g < c1 ? slab[r1 * n2 + e] : hsnap[(i64)hb * n2 + (g - c1) * n2 + e % n2];
            }
        }
        __syncthreads();
        if (n2 % (16 / (i64)sizeof(T)) == 0)
            lds_compute<T, true, 16 / (int)sizeof(T)>(slab, p, i0, r0, nr, hb, lds_rows, lds);
        else
            lds_compute<T, true, 1>(slab, p, i0, r0, nr, hb, lds_rows, lds);
        __syncthreads();  // every stencil read of this segment is done
        if (r1 < c1) {
            for (i64 e = threadIdx.x; e < (i64)hb * n2; e += nthr) lds[e] = lds[nr * n2 + e];
            __syncthreads();
        }
    }
}

// One step over rows [lo, hi) of slab i0 (unwrapped row numbers: rows
// before 0 and from n1 are the line's periodic continuation), reading src
// and writing dst, both indexed by row - base (LDS buffers of the staged
// rows) -- or, for a dst in HBM (TOHBM), by the row itself.  The stencil of
// row r lies delta = ipos1 - r rows from it (normalised into (-n1/2, n1/2]),
// the foot being that of row r mod n1.  Arithmetic of the storage type T,
// in FP64.  Thread (row, group of VC columns).
template <class T, int VC, bool TOHBM, class S, class D>
__device__ __forceinline__ void seg_step(const S *__restrict__ src, D *__restrict__ dst, const AdvDev &p, i64 i0,
                                         i64 lo, i64 hi, i64 base) {
    const i64 n1 = p.n1, n2 = p.n2;
    const int nthr = blockDim.x;
    const i64 nch = n2 / VC;
    const int cw = (int)min<i64>(nch, nthr);
    const int rows_per_pass = nthr / cw;
    const int t_row = threadIdx.x / cw, t_col = threadIdx.x - t_row * cw;
    if (t_row >= rows_per_pass) return;
    const int half = (int)(n1 / 2);
    double coef[6];
    unsigned long long cached = ~0ull;
    for (i64 r = lo + t_row; r < hi; r += rows_per_pass) {
        i64 rw = r % n1;
        if (rw < 0) rw += n1;
        double d_prev1;
        const int leftNode = adv_foot(p, (int)i0, (int)rw, d_prev1);
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, d_prev1);
        if (bits != cached) {
            adv_lag_basis(d_prev1, coef);
            cached = bits;
        }
        int delta = leftNode - 2 - (int)rw;
        if (delta < -half) delta += (int)n1;
        if (delta > half) delta -= (int)n1;
        const S *row0 = src + (r - base + delta) * n2;
        for (i64 ch = t_col; ch < nch; ch += cw) {
            double acc[VC];
#pragma unroll
            for (int v = 0; v < VC; ++v) acc[v] = 0.;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
#pragma unroll
                for (int v = 0; v < VC; ++v) {
                    const double f = (double)row0[k * n2 + ch * VC + v];
                    if constexpr (sizeof(T) == 8) acc[v] = acc[v] + coef[k] * f;
                    else acc[v] = fma(coef[k], f, acc[v]);
                }
            }
            D *out = dst + (TOHBM ? r : r - base) * n2 + ch * VC;
#pragma unroll
            for (int v = 0; v < VC; ++v) out[v] = (D)acc[v];
        }
    }
}

// line_chunk_unit advancing ns steps per pass.  A segment stages its rows
// with ns*hb old rows before and ns*ha after (carried from the previous
// segment, read from HBM, or from the chunk's snapshot, as line_chunk_unit
// does with one step's halo); step t computes the rows that steps t+1 .. ns
// still need, [r0 - (ns-t)*hb, r1 + (ns-t)*ha), into the FP64 buffers w0 and
// w1 in turn; step ns writes [r0, r1) to HBM.  Needs seg_rows >= ns*hb_max.
template <class T>
__device__ __forceinline__ void line_chunk_steps_unit(T *__restrict__ slab, const T *__restrict__ hsnap,
                                                      const AdvDev &p, i64 i0, i64 c0, i64 c1, i64 seg_rows,
                                                      i64 hb_max, i64 ha_max, int ns, T *__restrict__ st,
                                                      double *__restrict__ w0, double *__restrict__ w1) {
    const i64 n2 = p.n2;
    const int nthr = blockDim.x;
    int hb, ha;
    slab_halo(p, i0, hb_max, ha_max, hb, ha);
    const i64 HB = (i64)ns * hb, HA = (i64)ns * ha;
    for (i64 r0 = c0; r0 < c1; r0 += seg_rows) {
        const i64 r1 = min(c1, r0 + seg_rows);
        const i64 nr = r1 - r0;
        const i64 base = r0 - HB;  // global row of buffer row 0
        if (r0 == c0)
            for (i64 e = threadIdx.x; e < HB * n2; e += nthr) st[e] = hsnap[e];
        for (i64 e = threadIdx.x; e < nr * n2; e += nthr) st[HB * n2 + e] = slab[r0 * n2 + e];
        for (i64 e = threadIdx.x; e < HA * n2; e += nthr) {
            const i64 g = r1 + e / n2;  // global row
            st[(HB + nr) * n2 + e] = g < c1 ? slab[r1 * n2 + e] : hsnap[HB * n2 + (g - c1) * n2 + e % n2];
        }
        __syncthreads();
        auto step = [&](const auto *src, auto *dst, i64 lo, i64 hi, auto tohbm_) __attribute__((always_inline)) {
            constexpr bool TOHBM = decltype(tohbm_)::value;
            if (n2 % 4 == 0)
                seg_step<T, 4, TOHBM>(src, dst, p, i0, lo, hi, base);
            else if (n2 % 2 == 0)
                seg_step<T, 2, TOHBM>(src, dst, p, i0, lo, hi, base);
            else
                seg_step<T, 1, TOHBM>(src, dst, p, i0, lo, hi, base);
            __syncthreads();
        };
        const T *s0 = st;
        step(s0, w0, r0 - (ns - 1) * hb, r1 + (ns - 1) * ha, std::false_type{});
        for (int t = 2; t < ns; ++t) {
            const i64 lo = r0 - (i64)(ns - t) * hb, hi = r1 + (i64)(ns - t) * ha;
            if (t % 2 == 0)
                step((const double *)w0, w1, lo, hi, std::false_type{});
            else
                step((const double *)w1, w0, lo, hi, std::false_type{});
        }
        step((const double *)(ns % 2 == 0 ? w0 : w1), slab, r0, r1, std::true_type{});
        if (r1 < c1) {  // old rows [r1 - HB, r1) move to the front for the next segment
            for (i64 e = threadIdx.x; e < HB * n2; e += nthr) st[e] = st[nr * n2 + e];
            __syncthreads();
        }
    }
}

// NT threads per work-group; FUSED: ns > 1 steps per pass (a.ns), else one.
template <class T, int NT, bool FUSED>
__global__ __launch_bounds__(NT) void adv_line_chunk_kernel(T *__restrict__ data, const T *__restrict__ halo,
                                                            LdsArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // fused steps: the staged rows (storage type), then two FP64 buffers of as many rows
    const i64 rows = a.ns * (a.hb_max + a.ha_max) + a.seg_rows;
    T *st = reinterpret_cast<T *>(lds_raw);
    double *w0 = reinterpret_cast<double *>(lds_raw + ((rows * a.p.n2 * (i64)sizeof(T) + 15) & ~(i64)15));
    for (i64 unit = blockIdx.x; unit < a.p.n0 * a.nchunk; unit += gridDim.x) {
        const i64 i0 = unit / a.nchunk, ch = unit - i0 * a.nchunk;
        const i64 chunk_rows = a.seg_rows * a.chunk_segs;
        const i64 c0 = ch * chunk_rows;
        const i64 c1 = min(a.p.n1, c0 + chunk_rows);
        const T *hs = halo + unit * a.ns * (a.hb_max + a.ha_max) * a.p.n2;
        if constexpr (FUSED)
            line_chunk_steps_unit<T>(data + i0 * a.p.n1 * a.p.n2, hs, a.p, i0, c0, c1, a.seg_rows, a.hb_max,
                                     a.ha_max, (int)a.ns, st, w0, w0 + rows * a.p.n2);
        else
            line_chunk_unit<T>(data + i0 * a.p.n1 * a.p.n2, hs, a.p, i0, c0, c1, a.seg_rows, a.hb_max, a.ha_max,
                               reinterpret_cast<T *>(lds_raw));
        __syncthreads();  // the next unit restages the LDS
    }
}

}  // namespace bkma
