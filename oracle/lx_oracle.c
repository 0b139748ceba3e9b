/*
 * lx_oracle.c -- CPU oracle of the tiled Cholesky / SPD-solve / refinement
 * path (see lx_oracle.h for what each function restates).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline of the benchmark.  Nothing in the product links or loads it.
 *
 * Tile arithmetic goes through the host Fortran-77 LAPACK/BLAS interface
 * (dpotrf_, dtrsm_, dsyrk_, dgemm_ and the s* twins), exactly the calls the
 * reference's tile tasks make (src/blocked_blas/tasks_potrf.c:20-24,
 * tasks_trsm.c:22-25, tasks_syrk.c:19-21, task_gemm.c:55-58), linked against
 * a sequential BLAS so that parallelism comes only from the task graph.
 */
#include "lx_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* F77 LAPACK/BLAS (LP64). */
extern void dpotrf_(const char *, const int *, double *, const int *, int *);
extern void spotrf_(const char *, const int *, float *, const int *, int *);
extern void dtrsm_(const char *, const char *, const char *, const char *, const int *, const int *,
                   const double *, const double *, const int *, double *, const int *);
extern void strsm_(const char *, const char *, const char *, const char *, const int *, const int *,
                   const float *, const float *, const int *, float *, const int *);
extern void dsyrk_(const char *, const char *, const int *, const int *, const double *, const double *,
                   const int *, const double *, double *, const int *);
extern void ssyrk_(const char *, const char *, const int *, const int *, const float *, const float *,
                   const int *, const float *, float *, const int *);
extern void dgemm_(const char *, const char *, const int *, const int *, const int *, const double *,
                   const double *, const int *, const double *, const int *, const double *, double *,
                   const int *);
extern void sgemm_(const char *, const char *, const int *, const int *, const int *, const float *,
                   const float *, const int *, const float *, const int *, const float *, float *,
                   const int *);

/* Reference priorities: chol_rlmain.c:16-19 and chol_llmain.c:17-20. */
#define PRIO_POTRF 9999
#define PRIO_TRSM 9999
#define PRIO_SYRK 9
#define PRIO_GEMM 9
#define PRIO_LL_SYRK 9999
#define PRIO_LL_POTRF 9999
#define PRIO_LL_GEMM 9
#define PRIO_LL_TRSM 9

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* ---- double instantiation ---- */
#define T double
#define P lxo_d
#define POTRF dpotrf_
#define TRSM dtrsm_
#define SYRK dsyrk_
#define GEMM dgemm_
#include "lx_oracle_tiled.inc"
#undef T
#undef P
#undef POTRF
#undef TRSM
#undef SYRK
#undef GEMM

/* ---- float instantiation ---- */
#define T float
#define P lxo_s
#define POTRF spotrf_
#define TRSM strsm_
#define SYRK ssyrk_
#define GEMM sgemm_
#include "lx_oracle_tiled.inc"
#undef T
#undef P
#undef POTRF
#undef TRSM
#undef SYRK
#undef GEMM

/* ------------------------------------------------------------------ */
/* drand48 restated from its POSIX definition.                         */

typedef struct { uint64_t x; } rng48;
static void rng48_seed(rng48 *r, long s) { r->x = (((uint64_t)(uint32_t)s) << 16 | 0x330Eu) & 0xFFFFFFFFFFFFull; }
static double rng48_next(rng48 *r)
{
    r->x = (0x5DEECE66Dull * r->x + 0xBull) & 0xFFFFFFFFFFFFull;
    return ldexp((double)r->x, -48);
}

void lxo_drand48_fill(long seed, int64_t count, double *out)
{
    rng48 r;
    rng48_seed(&r, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = rng48_next(&r);
}

/* GENMAT_SPD, src/matutil/genmat.c:49-70: j-outer, i >= j, mirrored. */
void lxo_genmat_spd(int n, int lda, long seed, int diag_mode, double *A)
{
    rng48 r;
    rng48_seed(&r, seed);
    for (int j = 0; j < n; ++j) {
        for (int i = j; i < n; ++i) {
            double d = rng48_next(&r);
            if (i == j) d = diag_mode ? (double)n : d + n;
            A[(size_t)j * lda + i] = A[(size_t)i * lda + j] = d;
        }
    }
}

/* GENMAT, src/matutil/genmat.c:106-128. */
void lxo_genmat(int m, int n, int lda, long seed, double *A)
{
    rng48 r;
    rng48_seed(&r, seed);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) A[(size_t)j * lda + i] = rng48_next(&r);
}

/* GENMAT_SYM_FULL, tests/dense_tests/test_aux.h:114-128. */
void lxo_genmat_sym_full(int n, long seed, double *A)
{
    rng48 r;
    rng48_seed(&r, seed);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) {
            double d = rng48_next(&r);
            A[(size_t)j * n + i] = A[(size_t)i * n + j] = d;
        }
    for (int j = 0; j < n; ++j) A[(size_t)j * n + j] += 10.0 * n;
}

/* ------------------------------------------------------------------ */

int lxo_dchol_rl(int n, int b, double *A, int lda, int nthreads) { return lxo_dchol_rl_impl(n, b, A, lda, nthreads); }
int lxo_dchol_ll(int n, int b, double *A, int lda, int nthreads) { return lxo_dchol_ll_impl(n, b, A, lda, nthreads); }
int lxo_schol_rl(int n, int b, float *A, int lda, int nthreads) { return lxo_schol_rl_impl(n, b, A, lda, nthreads); }
int lxo_schol_ll(int n, int b, float *A, int lda, int nthreads) { return lxo_schol_ll_impl(n, b, A, lda, nthreads); }
int lxo_dchol_hrl(int mt, int b, double **Ah, int nthreads) { return lxo_dchol_hrl_impl(mt, b, Ah, nthreads); }
int lxo_dchol_hll(int mt, int b, double **Ah, int nthreads) { return lxo_dchol_hll_impl(mt, b, Ah, nthreads); }
int lxo_schol_hrl(int mt, int b, float **Ah, int nthreads) { return lxo_schol_hrl_impl(mt, b, Ah, nthreads); }
int lxo_schol_hll(int mt, int b, float **Ah, int nthreads) { return lxo_schol_hll_impl(mt, b, Ah, nthreads); }

void lxo_dltrsm_lower(int trans, int bm, int bn, int m, int n, double alpha,
                      const double *A, int lda, double *B, int ldb, int nthreads)
{
    lxo_dltrsm_lower_impl(trans, bm, bn, m, n, alpha, A, lda, B, ldb, nthreads);
}

void lxo_sltrsm_lower(int trans, int bm, int bn, int m, int n, float alpha,
                      const float *A, int lda, float *B, int ldb, int nthreads)
{
    lxo_sltrsm_lower_impl(trans, bm, bn, m, n, alpha, A, lda, B, ldb, nthreads);
}

/* dspdsolv_main, src/spdsolv/spdsolv_main.c:10-20. */
int lxo_dspdsolv(int m, int n, int b, double *A, double *B, int nthreads)
{
    int info = lxo_dchol_rl(m, b, A, m, nthreads);
    lxo_dltrsm_lower(0, b, n, m, n, 1.0, A, m, B, m, nthreads);
    lxo_dltrsm_lower(1, b, n, m, n, 1.0, A, m, B, m, nthreads);
    return info;
}

/* ------------------------------------------------------------------ */
/* Checks.                                                             */

double lxo_relerr_lower(int n, const double *A, int lda, const double *L, int ldl)
{
    /* norm1(tril(|A| - |L|)) / norm1(tril(|A|)), densutil.c:50-99 */
    double nd = 0.0, na = 0.0;
    for (int j = 0; j < n; ++j) {
        double sa = 0.0, sd = 0.0;
        for (int i = j; i < n; ++i) {
            double a = fabs(A[(size_t)j * lda + i]);
            sa += a;
            sd += fabs(a - fabs(L[(size_t)j * ldl + i]));
        }
        if (sd > nd) nd = sd;
        if (sa > na) na = sa;
    }
    return nd / na;
}

double lxo_potrf_resid(int n, const double *A0, int lda, const double *L, int ldl)
{
    /* ||L L^T - A0||_inf / (||A0||_inf N eps), potrf_test.c:6-37; L is the
     * lower triangle of L (upper part ignored). */
    double *R = (double *)calloc((size_t)n * n, sizeof(double));
    double *Lc = (double *)calloc((size_t)n * n, sizeof(double));
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) Lc[(size_t)j * n + i] = L[(size_t)j * ldl + i];
    const double one = 1.0, zero = 0.0;
    dgemm_("N", "T", &n, &n, &n, &one, Lc, &n, Lc, &n, &zero, R, &n);
    double rmax = 0.0, amax = 0.0;
    for (int i = 0; i < n; ++i) {
        double rs = 0.0, as = 0.0;
        for (int j = 0; j < n; ++j) {
            rs += fabs(R[(size_t)j * n + i] - A0[(size_t)j * lda + i]);
            as += fabs(A0[(size_t)j * lda + i]);
        }
        if (rs > rmax) rmax = rs;
        if (as > amax) amax = as;
    }
    free(R);
    free(Lc);
    return rmax / (amax * n * DBL_EPSILON / 2.0); /* LAPACK dlamch('e') = 2^-53 */
}

static double mat_inf_norm(int n, const double *A, int lda)
{
    double *rs = (double *)calloc(n, sizeof(double));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) rs[i] += fabs(A[(size_t)j * lda + i]);
    double m = 0.0;
    for (int i = 0; i < n; ++i)
        if (rs[i] > m) m = rs[i];
    free(rs);
    return m;
}

/* R = B - A X over the full stored A (FP64). */
static void residual(int n, int nrhs, const double *A, int lda, const double *X, int ldx,
                     const double *B, int ldb, double *R)
{
    for (int c = 0; c < nrhs; ++c) memcpy(&R[(size_t)c * n], &B[(size_t)c * ldb], sizeof(double) * n);
    const double one = 1.0, mone = -1.0;
    dgemm_("N", "N", &n, &nrhs, &n, &mone, A, &lda, X, &ldx, &one, R, &n);
}

static double col_inf(int n, const double *v)
{
    double m = 0.0;
    for (int i = 0; i < n; ++i)
        if (fabs(v[i]) > m || isnan(v[i])) m = isnan(v[i]) ? NAN : fabs(v[i]);
    return m;
}

double lxo_backward_error(int n, int nrhs, const double *A, int lda, const double *X, int ldx,
                          const double *B, int ldb)
{
    double *R = (double *)malloc(sizeof(double) * (size_t)n * nrhs);
    residual(n, nrhs, A, lda, X, ldx, B, ldb, R);
    double an = mat_inf_norm(n, A, lda), worst = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        double e = col_inf(n, &R[(size_t)c * n]) / (an * col_inf(n, &X[(size_t)c * ldx]));
        if (!(e <= worst)) worst = e;
    }
    free(R);
    return worst;
}

/* ------------------------------------------------------------------ */
/* Refinement loop: itref.c:999-1101's contract (x0 = solve(b); r = b - A x in
 * FP64; loop { test; d = solve(r); x += d; r = b - A x }) with the inner
 * solver replaced by the FP32 tiled Cholesky solve and the convergence test
 * made normwise-relative and taken before each correction. */
int lxo_dposv_ir(int n, int nrhs, int b, const double *A, int lda, const double *B, int ldb,
                 double *X, int ldx, const lxo_ir_opts *o, lxo_ir_status *st)
{
    lxo_ir_opts def = {30, 1e-14, 0.5, 0};
    if (!o) o = &def;
    memset(st, 0, sizeof(*st));
    int nt = o->nthreads;
    float *Af = (float *)malloc(sizeof(float) * (size_t)n * n);
    float *Sf = (float *)malloc(sizeof(float) * (size_t)n * nrhs);
    double *R = (double *)malloc(sizeof(double) * (size_t)n * nrhs);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Af[(size_t)j * n + i] = (float)A[(size_t)j * lda + i];
    int info = lxo_schol_rl(n, b, Af, n, nt);
    double anorm = mat_inf_norm(n, A, lda);
    int fallback = info != 0;
    double prev = INFINITY, berr = INFINITY;
    if (!fallback) {
        for (int c = 0; c < nrhs; ++c)
            for (int i = 0; i < n; ++i) Sf[(size_t)c * n + i] = (float)B[(size_t)c * ldb + i];
        lxo_sltrsm_lower(0, b, nrhs, n, nrhs, 1.0f, Af, n, Sf, n, nt);
        lxo_sltrsm_lower(1, b, nrhs, n, nrhs, 1.0f, Af, n, Sf, n, nt);
        for (int c = 0; c < nrhs; ++c)
            for (int i = 0; i < n; ++i) X[(size_t)c * ldx + i] = Sf[(size_t)c * n + i];
        for (int it = 0;; ++it) {
            residual(n, nrhs, A, lda, X, ldx, B, ldb, R);
            double worst = 0.0, rn = 0.0;
            for (int c = 0; c < nrhs; ++c) {
                double r = col_inf(n, &R[(size_t)c * n]);
                double e = r / (anorm * col_inf(n, &X[(size_t)c * ldx]));
                if (!(e <= worst)) worst = e;
                if (!(r <= rn)) rn = r;
            }
            berr = worst;
            if (berr <= o->tol) break;
            if (it >= o->max_sweeps || !(rn <= o->stall_ratio * prev) || isnan(rn)) {
                fallback = 1;
                break;
            }
            prev = rn;
            for (int c = 0; c < nrhs; ++c)
                for (int i = 0; i < n; ++i) Sf[(size_t)c * n + i] = (float)R[(size_t)c * n + i];
            lxo_sltrsm_lower(0, b, nrhs, n, nrhs, 1.0f, Af, n, Sf, n, nt);
            lxo_sltrsm_lower(1, b, nrhs, n, nrhs, 1.0f, Af, n, Sf, n, nt);
            for (int c = 0; c < nrhs; ++c)
                for (int i = 0; i < n; ++i) X[(size_t)c * ldx + i] += (double)Sf[(size_t)c * n + i];
            st->sweeps = it + 1;
        }
    }
    if (fallback) {
        double *Ad = (double *)malloc(sizeof(double) * (size_t)n * n);
        for (int j = 0; j < n; ++j) memcpy(&Ad[(size_t)j * n], &A[(size_t)j * lda], sizeof(double) * n);
        info = lxo_dchol_rl(n, b, Ad, n, nt);
        for (int c = 0; c < nrhs; ++c) memcpy(&X[(size_t)c * ldx], &B[(size_t)c * ldb], sizeof(double) * n);
        lxo_dltrsm_lower(0, b, nrhs, n, nrhs, 1.0, Ad, n, X, ldx, nt);
        lxo_dltrsm_lower(1, b, nrhs, n, nrhs, 1.0, Ad, n, X, ldx, nt);
        free(Ad);
        berr = lxo_backward_error(n, nrhs, A, lda, X, ldx, B, ldb);
    }
    st->fallback = fallback;
    st->info = info;
    st->berr = berr;
    free(Af);
    free(Sf);
    free(R);
    return info;
}

double lxo_wtime(void) { return omp_get_wtime(); }
int lxo_max_threads(void) { return omp_get_max_threads(); }
