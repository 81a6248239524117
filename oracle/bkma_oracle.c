/*
 * bkma_oracle.c -- CPU restatement of the reference's BKMA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (libbkma_amd.so and the C++ host code) never links it.
 *
 * It restates, in plain C99 + OpenMP, with -ffp-contract=off so that every
 * multiply and add rounds separately exactly as written in the reference:
 *
 *   - ADVParams::update_deltas          src/config/AdvectionParams.cpp:96-103
 *   - AdvectionSolver::coord            src/solvers/AdvectionSolver.hpp:23-26
 *   - AdvectionSolver::lag_basis        src/solvers/AdvectionSolver.hpp:31-51
 *   - AdvectionSolver::displ            src/solvers/AdvectionSolver.hpp:56-66
 *   - AdvectionSolver::operator()       src/solvers/AdvectionSolver.hpp:71-98
 *   - AdaptiveWg line semantics (compute every window cell of the line into
 *     scratch from the old line, then copy back in place)
 *                                       src/core/AdaptiveWg.hpp:119-134
 *   - ConvSolver::operator()            src/solvers/ConvSolver.hpp:20-44
 *   - fill_buffer_adv                   src/tools/init.hpp:31-46
 *   - validate_result_adv (analytic L1) src/tools/validation.hpp:9-80
 *
 * plus two restatements that have no reference counterpart and pin the
 * MI355X-specific kernels:
 *   - orc_advect_f32_mixed: FP32 storage, FP64 foot/index/coefficients and
 *     FP64 fused accumulation (the arithmetic of the tuned FP32 kernel);
 *   - orc_advect_f32_blocked: the same with several steps per pass kept in
 *     FP64 and one rounding to FP32 per pass;
 *   - orc_conv_torch_f64:   textbook Conv1d (torch.nn.functional.conv1d
 *     semantics), the contraction the MFMA kernels implement.
 *
 * Pinning status of each function (what checks it besides itself) is stated
 * in tests/test_oracle.py; every function here is pinned by at least one
 * check that does not restate its formula.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Field-for-field the same layout as bkma_adv_params in include/bkma_amd.h,
 * which mirrors ADVParams (src/config/AdvectionParams.hpp:10-57). */
typedef struct {
    uint64_t n0, n1, n2;
    uint64_t maxIter;
    uint64_t pref_wg_size;
    uint64_t seq_size0, seq_size2;
    double dt, dx, dvx, inv_dx;
    double minRealX, maxRealX, realWidthX;
    double minRealVx, maxRealVx;
    float percent_loc;
    int32_t gpu;
    int32_t outputSolution;
    int32_t _pad;
} orc_adv_params;

int orc_abi_sizeof_params(void) { return (int)sizeof(orc_adv_params); }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Counter-based generator shared with the GPU fill kernels.           */
/* ------------------------------------------------------------------ */
static inline uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* U[0,1) with 53 random bits, keyed by (seed, idx). */
double orc_uniform(uint64_t seed, uint64_t idx) {
    uint64_t key = orc_mix64(seed ^ 0x5851F42D4C957F2DULL);
    uint64_t z = orc_mix64(key + (idx + 1ULL) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 0x1.0p-53;
}

void orc_fill_uniform_f64(double *out, uint64_t n, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = orc_uniform(seed, (uint64_t)i);
}

void orc_fill_uniform_f32(float *out, uint64_t n, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = (float)orc_uniform(seed, (uint64_t)i);
}

/* ------------------------------------------------------------------ */
/* Parameters                                                          */
/* ------------------------------------------------------------------ */
/* src/config/AdvectionParams.cpp:96-103 */
void orc_update_deltas(orc_adv_params *p) {
    p->realWidthX = p->maxRealX - p->minRealX;
    p->dx = p->realWidthX / (double)p->n1;
    p->dvx = (p->maxRealVx - p->minRealVx) / (double)p->n0;
    p->inv_dx = 1 / p->dx;
}

/* ------------------------------------------------------------------ */
/* AdvectionSolver                                                     */
/* ------------------------------------------------------------------ */
/* src/solvers/AdvectionSolver.hpp:10-11 */
static const double LOC[6] = {-1. / 24, 1. / 24., -1. / 12., 1. / 12., -1. / 24., 1. / 24.};

/* src/solvers/AdvectionSolver.hpp:23-26 */
static inline double coord(int i, double minValue, double delta) { return minValue + i * delta; }

/* src/solvers/AdvectionSolver.hpp:31-51 (products evaluated left to right) */
void orc_lag_basis(double px, double *coef) {
    const double pxm2 = px - 2.;
    const double sqrpxm2 = pxm2 * pxm2;
    const double pxm2_01 = pxm2 * (pxm2 - 1.);

    coef[0] = LOC[0] * pxm2_01 * (pxm2 + 1.) * (pxm2 - 2.) * (pxm2 - 1.);
    coef[1] = LOC[1] * pxm2_01 * (pxm2 - 2.) * (5 * sqrpxm2 + pxm2 - 8.);
    coef[2] = LOC[2] * (pxm2 - 1.) * (pxm2 - 2.) * (pxm2 + 1.) * (5 * sqrpxm2 - 3 * pxm2 - 6.);
    coef[3] = LOC[3] * pxm2 * (pxm2 + 1.) * (pxm2 - 2.) * (5 * sqrpxm2 - 7 * pxm2 - 4.);
    coef[4] = LOC[4] * pxm2_01 * (pxm2 + 1.) * (5 * sqrpxm2 - 11 * pxm2 - 2.);
    coef[5] = LOC[5] * pxm2_01 * pxm2 * (pxm2 + 1.) * (pxm2 - 2.);
}

/* src/solvers/AdvectionSolver.hpp:56-66 */
double orc_displ(const orc_adv_params *p, int i1, int i0) {
    const double x = coord(i1, p->minRealX, p->dx);
    const double vx = coord(i0, p->minRealVx, p->dvx);
    const double displx = p->dt * vx;
    return p->minRealX + fmod(p->realWidthX + x - displx - p->minRealX, p->realWidthX);
}

/* Foot location of cell (i0,i1): returns leftNode, writes d_prev1.
 * src/solvers/AdvectionSolver.hpp:76-85 */
int orc_adv_foot(const orc_adv_params *p, int i0, int i1, double *d_prev1) {
    const double xFoot = orc_displ(p, i1, i0);
    const int leftNode = (int)floor((xFoot - p->minRealX) * p->inv_dx);
    *d_prev1 = 2 + p->inv_dx * (xFoot - coord(leftNode, p->minRealX, p->dx));
    return leftNode;
}

/* src/solvers/AdvectionSolver.hpp:71-98.  `line` points at element i1=0 of
 * the (i0, :, i2) line, whose elements are `stride` apart. */
double orc_adv_cell(const orc_adv_params *p, const double *line, int64_t stride, int i0, int i1) {
    double d_prev1;
    const int leftNode = orc_adv_foot(p, i0, i1, &d_prev1);
    double coef[6];
    orc_lag_basis(d_prev1, coef);
    const int ipos1 = leftNode - 2;
    double value = 0.;
    for (int k = 0; k <= 5; k++) {
        /* (size_t)n1 + ipos1 + k is unsigned in the reference; for every
         * reachable ipos1 (>= -2) the sum is non-negative, so this is the
         * mathematical modulus. */
        const uint64_t id1 = (p->n1 + (uint64_t)(int64_t)ipos1 + (uint64_t)k) % p->n1;
        value += coef[k] * line[(int64_t)id1 * stride];
    }
    return value;
}

/* One in-place AdaptiveWg step on FP64 data [n0][n1][n2]:
 * every line (i0,:,i2) gets out[i1] = solver(old line, i0, i1, i2).
 * src/core/AdaptiveWg.hpp:119-134 with window() == 1. */
static void adv_step_f64(double *data, const orc_adv_params *p) {
    const int64_t n0 = (int64_t)p->n0, n1 = (int64_t)p->n1, n2 = (int64_t)p->n2;
#pragma omp parallel
    {
        double *old = (double *)malloc(sizeof(double) * (size_t)n1);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < n0 * n2; ++line) {
            const int64_t i0 = line / n2, i2 = line % n2;
            double *base = data + (i0 * n1) * n2 + i2;
            for (int64_t i1 = 0; i1 < n1; ++i1) old[i1] = base[i1 * n2];
            for (int64_t i1 = 0; i1 < n1; ++i1)
                base[i1 * n2] = orc_adv_cell(p, old, 1, (int)i0, (int)i1);
        }
        free(old);
    }
}

void orc_advect_f64(double *data, const orc_adv_params *p, uint64_t nsteps) {
    for (uint64_t t = 0; t < nsteps; ++t) adv_step_f64(data, p);
}

/* FP32 storage, FP64 foot / index / lag_basis exactly as the reference,
 * FP64 fused accumulation acc = fma(coef[k], (double)f[k], acc) for
 * k = 0..5 from acc = 0, one rounding to FP32 at the store. */
static void adv_step_f32_mixed(float *data, const orc_adv_params *p) {
    const int64_t n0 = (int64_t)p->n0, n1 = (int64_t)p->n1, n2 = (int64_t)p->n2;
#pragma omp parallel
    {
        float *old = (float *)malloc(sizeof(float) * (size_t)n1);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < n0 * n2; ++line) {
            const int64_t i0 = line / n2, i2 = line % n2;
            float *base = data + (i0 * n1) * n2 + i2;
            for (int64_t i1 = 0; i1 < n1; ++i1) old[i1] = base[i1 * n2];
            for (int64_t i1 = 0; i1 < n1; ++i1) {
                double d_prev1;
                const int leftNode = orc_adv_foot(p, (int)i0, (int)i1, &d_prev1);
                double coef[6];
                orc_lag_basis(d_prev1, coef);
                const int ipos1 = leftNode - 2;
                double acc = 0.;
                for (int k = 0; k <= 5; k++) {
                    const uint64_t id1 = (p->n1 + (uint64_t)(int64_t)ipos1 + (uint64_t)k) % p->n1;
                    acc = fma(coef[k], (double)old[id1], acc);
                }
                base[i1 * n2] = (float)acc;
            }
        }
        free(old);
    }
}

void orc_advect_f32_mixed(float *data, const orc_adv_params *p, uint64_t nsteps) {
    for (uint64_t t = 0; t < nsteps; ++t) adv_step_f32_mixed(data, p);
}

/* One FP64 step of a single line with the FP32-storage arithmetic above
 * (fused accumulation), out-of-place: dst[i1] from src. */
static void adv_line_fma_f64(const orc_adv_params *p, int i0, const double *src, double *dst) {
    for (int64_t i1 = 0; i1 < (int64_t)p->n1; ++i1) {
        double d_prev1;
        const int leftNode = orc_adv_foot(p, i0, (int)i1, &d_prev1);
        double coef[6];
        orc_lag_basis(d_prev1, coef);
        const int ipos1 = leftNode - 2;
        double acc = 0.;
        for (int k = 0; k <= 5; k++) {
            const uint64_t id1 = (p->n1 + (uint64_t)(int64_t)ipos1 + (uint64_t)k) % p->n1;
            acc = fma(coef[k], src[id1], acc);
        }
        dst[i1] = acc;
    }
}

/* FP32 storage with `ns` steps per pass: each pass widens the line to FP64,
 * makes its steps with the fused FP64 arithmetic of orc_advect_f32_mixed
 * without rounding in between, and rounds to FP32 once.  nsteps / ns full
 * passes, then one pass of nsteps % ns steps.  ns = 1 is
 * orc_advect_f32_mixed. */
void orc_advect_f32_blocked(float *data, const orc_adv_params *p, uint64_t nsteps, uint64_t ns) {
    if (ns == 0) ns = 1;
    const int64_t n0 = (int64_t)p->n0, n1 = (int64_t)p->n1, n2 = (int64_t)p->n2;
#pragma omp parallel
    {
        double *a = (double *)malloc(sizeof(double) * (size_t)n1);
        double *b = (double *)malloc(sizeof(double) * (size_t)n1);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < n0 * n2; ++line) {
            const int64_t i0 = line / n2, i2 = line % n2;
            float *base = data + (i0 * n1) * n2 + i2;
            for (uint64_t done = 0; done < nsteps;) {
                const uint64_t k = nsteps - done < ns ? nsteps - done : ns;
                for (int64_t i1 = 0; i1 < n1; ++i1) a[i1] = (double)base[i1 * n2];
                for (uint64_t s = 0; s < k; ++s) {
                    adv_line_fma_f64(p, (int)i0, a, b);
                    double *t = a;
                    a = b;
                    b = t;
                }
                for (int64_t i1 = 0; i1 < n1; ++i1) base[i1 * n2] = (float)a[i1];
                done += k;
            }
        }
        free(a);
        free(b);
    }
}

/* FP32 storage through the plain executors: the reference's FP64
 * AdvectionSolver (separately rounded multiply and add) on values widened
 * from FP32, one rounding to FP32 at the store. */
static void adv_step_f32_plain(float *data, const orc_adv_params *p) {
    const int64_t n0 = (int64_t)p->n0, n1 = (int64_t)p->n1, n2 = (int64_t)p->n2;
#pragma omp parallel
    {
        double *old = (double *)malloc(sizeof(double) * (size_t)n1);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < n0 * n2; ++line) {
            const int64_t i0 = line / n2, i2 = line % n2;
            float *base = data + (i0 * n1) * n2 + i2;
            for (int64_t i1 = 0; i1 < n1; ++i1) old[i1] = (double)base[i1 * n2];
            for (int64_t i1 = 0; i1 < n1; ++i1) base[i1 * n2] = (float)orc_adv_cell(p, old, 1, (int)i0, (int)i1);
        }
        free(old);
    }
}

void orc_advect_f32_plain(float *data, const orc_adv_params *p, uint64_t nsteps) {
    for (uint64_t t = 0; t < nsteps; ++t) adv_step_f32_plain(data, p);
}

/* src/tools/init.hpp:31-46: data(i0,i1,i2) = sin(4 * x * M_PI). */
void orc_fill_adv_sin4(double *data, const orc_adv_params *p) {
    const int64_t n0 = (int64_t)p->n0, n1 = (int64_t)p->n1, n2 = (int64_t)p->n2;
#pragma omp parallel for schedule(static)
    for (int64_t i0 = 0; i0 < n0; ++i0)
        for (int64_t i1 = 0; i1 < n1; ++i1) {
            const double x = p->minRealX + (double)i1 * p->dx;
            const double v = sin(4 * x * M_PI);
            for (int64_t i2 = 0; i2 < n2; ++i2) data[(i0 * n1 + i1) * n2 + i2] = v;
        }
}

/* src/tools/validation.hpp:9-80: per i2 slab, mean |f - sin(4 pi (x - v t))|
 * over (i0,i1) with t = maxIter * dt; returns the max over i2 (and the min
 * through *lowest when non-null). */
double orc_l1_error_adv(const double *data, const orc_adv_params *p, double *lowest) {
    const int64_t n0 = (int64_t)p->n0, n1 = (int64_t)p->n1, n2 = (int64_t)p->n2;
    double hi = -1.0, lo = INFINITY;
    for (int64_t i2 = 0; i2 < n2; ++i2) {
        double err = 0.0;
#pragma omp parallel for reduction(+ : err) schedule(static)
        for (int64_t i0 = 0; i0 < n0; ++i0)
            for (int64_t i1 = 0; i1 < n1; ++i1) {
                const double f = data[(i0 * n1 + i1) * n2 + i2];
                const double x = p->minRealX + (double)i1 * p->dx;
                const double v = p->minRealVx + (double)i0 * p->dvx;
                const double t = (double)p->maxIter * p->dt;
                err += fabs(f - sin(4 * M_PI * (x - v * t)));
            }
        const double l1 = err / (double)(n1 * n0);
        if (l1 > hi) hi = l1;
        if (l1 < lo) lo = l1;
    }
    if (lowest) *lowest = lo;
    return hi;
}

/* ------------------------------------------------------------------ */
/* ConvSolver (reference arithmetic, through the AdaptiveWg window)     */
/* ------------------------------------------------------------------ */
/* One in-place step of bkma_run<ConvSolver, AdaptiveWg> on [n0][n1][n2]
 * with n1 = C * L and w laid out (K, C_in, C_out):
 *   for ii1 in [K-1, n1): scratch[ii1-(K-1)] = solver(old line, ii1)
 *   for iw  in [0, n1-K+1): line[iw] = scratch[iw]
 * The solver (src/solvers/ConvSolver.hpp:20-44): oc = ii1 / L,
 * il = ii1 - oc*L, sum = b[oc]; for ic, for k: if il - k >= 0:
 * sum += line[il - k] * w(k, ic, oc).  Cells ii1 < K-1 have a negative
 * window index; the reference's unsigned index makes those writes land
 * outside the line's scratch (undefined), so they are not performed here
 * and positions [n1-K+1, n1) keep their old values. */
void orc_conv_literal_f64(double *data, uint64_t n0, uint64_t n1, uint64_t n2, const double *w, const double *b,
                          uint64_t K, uint64_t C, uint64_t L) {
    const int64_t nw = (int64_t)n1 - ((int64_t)K - 1);
#pragma omp parallel
    {
        double *old = (double *)malloc(sizeof(double) * (size_t)n1);
        double *scr = (double *)malloc(sizeof(double) * (size_t)(nw > 0 ? nw : 1));
#pragma omp for schedule(static)
        for (int64_t line = 0; line < (int64_t)(n0 * n2); ++line) {
            const int64_t i0 = line / (int64_t)n2, i2 = line % (int64_t)n2;
            double *base = data + (i0 * (int64_t)n1) * (int64_t)n2 + i2;
            for (int64_t i1 = 0; i1 < (int64_t)n1; ++i1) old[i1] = base[i1 * (int64_t)n2];
            for (int64_t ii1 = (int64_t)K - 1; ii1 < (int64_t)n1; ++ii1) {
                const uint64_t oc = (uint64_t)ii1 / L;
                const uint64_t il = (uint64_t)ii1 - oc * L;
                double sum = b[oc];
                for (uint64_t ic = 0; ic < C; ++ic)
                    for (uint64_t k = 0; k < K; ++k) {
                        const int input_idx = (int)(il - k);
                        if ((uint64_t)(int64_t)input_idx < n1) sum += old[input_idx] * w[(k * C + ic) * C + oc];
                    }
                scr[ii1 - ((int64_t)K - 1)] = sum;
            }
            for (int64_t iw = 0; iw < nw; ++iw) base[iw * (int64_t)n2] = scr[iw];
        }
        free(old);
        free(scr);
    }
}

/* Textbook Conv1d, stride 1, no padding, written in place:
 * x[B][C][L], w[C_out = C][C_in = C][K], b[C];
 * y[b][oc][j] = b[oc] + sum_ic sum_k w[oc][ic][k] * x[b][ic][j+k],
 * j in [0, L-K+1) overwrites x[b][oc][j]; x[b][oc][L-K+1 .. L) is kept.
 * Accumulation order: ic outer, k inner, starting from the bias. */
void orc_conv_torch_f64(double *x, uint64_t B, uint64_t C, uint64_t L, const double *w, const double *b, uint64_t K) {
    const int64_t Lout = (int64_t)L - (int64_t)K + 1;
    if (Lout <= 0) return;
#pragma omp parallel
    {
        double *y = (double *)malloc(sizeof(double) * (size_t)(C * (uint64_t)Lout));
#pragma omp for schedule(static)
        for (int64_t bb = 0; bb < (int64_t)B; ++bb) {
            double *xb = x + (uint64_t)bb * C * L;
            for (uint64_t oc = 0; oc < C; ++oc)
                for (int64_t j = 0; j < Lout; ++j) {
                    double s = b[oc];
                    for (uint64_t ic = 0; ic < C; ++ic) {
                        const double *wr = w + (oc * C + ic) * K;
                        const double *xr = xb + ic * L + j;
                        for (uint64_t k = 0; k < K; ++k) s += wr[k] * xr[k];
                    }
                    y[oc * (uint64_t)Lout + (uint64_t)j] = s;
                }
            for (uint64_t oc = 0; oc < C; ++oc)
                memcpy(xb + oc * L, y + oc * (uint64_t)Lout, sizeof(double) * (size_t)Lout);
        }
        free(y);
    }
}
