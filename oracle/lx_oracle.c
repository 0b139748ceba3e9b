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
        const double r = col_inf(n, &R[(size_t)c * n]);
        /* a zero residual is an exact solution: backward error 0 (b_c = 0, x_c = 0
         * would otherwise give 0 / 0) */
        double e = r == 0.0 ? 0.0 : r / (an * col_inf(n, &X[(size_t)c * ldx]));
        if (!(e <= worst)) worst = e;
    }
    free(R);
    return worst;
}

/* ------------------------------------------------------------------ */
/* GENMAT_COVARIANT (genmat.c:263-302): column j holds |j - i|^-p above the
 * diagonal, 1 + (j + 1)^(1/k) on it, |i - j|^-p below, in storage order. */
int lxo_genmat_covariant(int n, double p, double k, double *A)
{
    if (p <= 0.0) return -1;
    const double mp = -p, kinv = 1.0 / k;
    for (int j = 0; j < n; ++j) {
        double *ap = &A[(size_t)j * n];
        for (int i = 0; i < j; ++i) *ap++ = pow((double)(j - i), mp);
        *ap++ = 1.0 + pow((double)(j + 1), kinv);
        for (int i = j + 1; i < n; ++i) *ap++ = pow((double)(i - j), mp);
    }
    return 0;
}

/* GENMAT_COSTAS (genmat.c:199-220): one draw per element, storage order. */
void lxo_genmat_costas(int m, int n, long seed, double *A)
{
    srand48(seed);
    const size_t s = (size_t)m * n;
    for (size_t c = 0; c < s; ++c) A[c] = drand48() > 0.5 ? 1.0 : -1.0;
}

/* dgenmat_config (genmat_config.c:34-133): the switch on ompssmat_t with the
 * default parameters of :37 when no configuration file is read. */
int lxo_dgenmat_config(int kind, int m, int n, int ldim, const double *pars, long seed, double *A)
{
    const double def[4] = {2.0, 2.0, 2.0, 2.0};
    if (!pars) pars = def;
    switch (kind) {
    case 0: /* GENMATCONF_SPD: GENMAT_SPD(m, ldim) */
        lxo_genmat_spd(m, ldim, seed, 0, A);
        return 0;
    case 1: /* GENMATCONF_COVAR: GENMAT_COVARIANT(m, pars[0], pars[1]) */
        return lxo_genmat_covariant(m, pars[0], pars[1], A);
    case 2: /* GENMATCONF_RAND: GENMAT(m, n, ldim) */
        lxo_genmat(m, n, ldim, seed, A);
        return 0;
    case 3: /* GENMATCONF_EYE: GENMAT_EYE(m, ldim) */
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < ldim; ++i) A[(size_t)j * ldim + i] = i == j ? 1.0 : 0.0;
        return 0;
    case 6: /* GENMATCONF_ZERO: GENMAT_ZERO(m, ldim) */
        memset(A, 0, sizeof(double) * (size_t)m * ldim);
        return 0;
    case 7: /* GENMATCONF_COSTAS: GENMAT_COSTAS(m, n) */
        lxo_genmat_costas(m, n, seed, A);
        return 0;
    default: /* DIAGDOM, ONES and unknown kinds: the reference returns NULL */
        return -1;
    }
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11): ten rounds; each round multiplies
 * ctr[0] and ctr[2] by the round multipliers, swaps the halves across and
 * xors in the key, whose words are bumped by the Weyl constants between
 * rounds. */
void lxo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double lxo_u53(uint32_t hi, uint32_t lo) { return (double)((((uint64_t)hi) << 21) ^ (lo >> 11)) * 0x1p-53; }

void lxo_gen_uniform(int m, int n, uint32_t seed, double lo, double hi, double *X, int ldx)
{
    const uint32_t key[2] = {seed, 0x89abcdefu};
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = 0; e < tot; ++e) {
        const uint32_t ctr[4] = {(uint32_t)e, (uint32_t)(e >> 32), 1u, 0u};
        uint32_t r[4];
        lxo_philox4x32_10(ctr, key, r);
        /* one rounding: the device evaluates lo + (hi - lo) u as a fused multiply-add */
        X[(e / m) * ldx + e % m] = fma(hi - lo, lxo_u53(r[0], r[1]), lo);
    }
}

void lxo_gen_gauss(int m, int n, uint32_t seed, double *G, int ldg)
{
    const uint32_t key[2] = {seed, 0x1234567u};
    const int64_t tot = (int64_t)m * n;
    for (int64_t e2 = 0; e2 < tot; e2 += 2) {
        const uint64_t q = (uint64_t)(e2 >> 1);
        const uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), 0u, 0u};
        uint32_t r[4];
        lxo_philox4x32_10(ctr, key, r);
        const double u0 = lxo_u53(r[0], r[1]) + 0x1p-53, u1 = lxo_u53(r[2], r[3]) + 0x1p-53;
        const double rad = sqrt(-2.0 * log(u0)), th = 6.283185307179586 * u1;
        G[(e2 / m) * ldg + e2 % m] = rad * cos(th);
        if (e2 + 1 < tot) G[((e2 + 1) / m) * ldg + (e2 + 1) % m] = rad * sin(th);
    }
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
                double e = r == 0.0 ? 0.0 : r / (anorm * col_inf(n, &X[(size_t)c * ldx]));
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
