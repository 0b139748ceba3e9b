/*
 * lx_oracle.h -- CPU oracle of the tiled Cholesky / SPD-solve / refinement path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (linompss_amd/liblinompss.so) never links, loads or calls it.
 *
 * It restates, in plain C over host LAPACK/BLAS (one sequential BLAS call per
 * tile task, as the reference does), the reference algorithms:
 *   - drand48 generators         src/matutil/genmat.c:49-70, 106-128,
 *                                199-220; tests/dense_tests/test_aux.h:114-128
 *   - covariance generator       src/matutil/genmat.c:263-302
 *   - generator dispatch         src/matutil/genmat_config.c:34-133
 *   - the benchmark's Philox4x32-10 streams (no reference counterpart;
 *     Salmon et al., SC'11; pinned by Random123's published known answers)
 *   - right-looking CHOL_RL      src/cholesky/dense/chol_rlmain.c:22-53
 *   - left-looking  CHOL_LL      src/cholesky/dense/chol_llmain.c:23-57
 *   - hypermatrix CHOL_HRL/HLL   src/cholesky/dense/chol_hrlmain.c:16-37,
 *                                chol_hllmain.c:17-45
 *   - blocked lower TRSM sweeps  src/blocked_blas/bblas_trsm.h:116-160
 *   - dspdsolv_main              src/spdsolv/spdsolv_main.c:10-20
 *   - refinement loop contract   src/itref/itref.c:999-1101 (inner solver is
 *                                the FP32 Cholesky solve, not CG)
 *   - checks                     src/matutil/densutil.c:50-99,
 *                                tests/dense_tests/potrf_test.c:6-37
 * The OmpSs task pragmas become OpenMP 4.5 `task depend(...) priority(...)`
 * over the first element of each tile, which is how the reference's tasks
 * are ordered by the Nanos++ runtime.
 */
#ifndef LX_ORACLE_H
#define LX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* POSIX drand48: X <- (0x5DEECE66D X + 0xB) mod 2^48, srand48(s): X = (s<<16)|0x330E */
void   lxo_drand48_fill(long seed, int64_t count, double *out);

/* GENMAT_SPD (genmat.c:49-70) with an explicit seed.  diag_mode 0: diagonal is
 * drand48()+n as in the reference; 1: the diagonal draw is consumed and the
 * diagonal set to n (the plain-FP64 workload's definition). Column-major. */
void   lxo_genmat_spd(int n, int lda, long seed, int diag_mode, double *A);
/* GENMAT (genmat.c:106-128): m x n uniform [0,1) column-major */
void   lxo_genmat(int m, int n, int lda, long seed, double *A);
/* GENMAT_SYM_FULL of the unit tests (test_aux.h:114-128): symmetric U[0,1) + 10n I */
void   lxo_genmat_sym_full(int n, long seed, double *A);

/* GENMAT_COVARIANT (genmat.c:263-302): n x n, lda = n, a_ij = |i - j|^-p off
 * the diagonal, a_jj = 1 + (j + 1)^(1/k).  Returns -1 (and writes nothing)
 * for p <= 0, as the reference returns NULL. */
int    lxo_genmat_covariant(int n, double p, double k, double *A);
/* GENMAT_COSTAS (genmat.c:199-220): m x n (ld m), each drand48() draw in
 * element order mapped to +1 when > 0.5, else -1; explicit seed. */
void   lxo_genmat_costas(int m, int n, long seed, double *A);
/* dgenmat_config (genmat_config.c:34-133) without the libconfig file (absent
 * here; the reference also skips it when built without libconfig): kind is
 * the ompssmat_t value (ompss_mat.h:5-14: 0 SPD, 1 COVAR, 2 RAND, 3 EYE,
 * 4 DIAGDOM, 5 ONES, 6 ZERO, 7 COSTAS); pars = the config's parameters
 * (NULL: the defaults {2, 2, 2, 2}); seed replaces time(NULL).  Writes into A
 * (m x n, ld ldim; COVAR and COSTAS use ld m as the reference allocates them)
 * and returns 0, or -1 where the reference returns NULL (DIAGDOM, ONES,
 * unknown kinds, COVAR with p <= 0).  EYE writes zeros off the diagonal
 * (the reference leaves them uninitialised, genmat.c:147-158). */
int    lxo_dgenmat_config(int kind, int m, int n, int ldim, const double *pars, long seed, double *A);

/* Workload generators of the benchmark (no counterpart in the reference):
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11), ctr[4] and key[2] in, out[4]. */
void   lxo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* X (m x n, ld ldx) uniform in [lo, hi): element e = j m + i from block
 * {e_lo, e_hi, 1, 0} under key {seed, 0x89abcdef}, u = ((x << 21) ^ (y >> 11)) 2^-53,
 * X = fma(hi - lo, u, lo) (one rounding). */
void   lxo_gen_uniform(int m, int n, uint32_t seed, double lo, double hi, double *X, int ldx);
/* G (m x n, ld ldg) N(0,1): elements 2q and 2q+1 (order e = j m + i) from
 * block {q_lo, q_hi, 0, 0} under key {seed, 0x1234567}: u0, u1 in (0, 1]
 * from (x, y) and (z, w) as above plus 2^-53, then Box-Muller
 * sqrt(-2 ln u0) (cos, sin)(2 pi u1) with the host libm. */
void   lxo_gen_gauss(int m, int n, uint32_t seed, double *G, int ldg);

/* Tiled Cholesky, lower, in place.  Returns 0 or the LAPACK-style info of the
 * first failing global column (1-based).  nthreads <= 0 means all cores. */
int    lxo_dchol_rl(int n, int b, double *A, int lda, int nthreads);
int    lxo_dchol_ll(int n, int b, double *A, int lda, int nthreads);
int    lxo_schol_rl(int n, int b, float *A, int lda, int nthreads);
int    lxo_schol_ll(int n, int b, float *A, int lda, int nthreads);
/* The hypermatrix variants CHOL_HRL / CHOL_HLL (chol_hrlmain.c:16-37,
 * chol_hllmain.c:17-45): tile (i, j) is Ah[j * mt + i], b x b with leading
 * dimension b, tiles i >= j read.  Same return convention (info = global column). */
int    lxo_dchol_hrl(int mt, int b, double **Ah, int nthreads);
int    lxo_dchol_hll(int mt, int b, double **Ah, int nthreads);
int    lxo_schol_hrl(int mt, int b, float **Ah, int nthreads);
int    lxo_schol_hll(int mt, int b, float **Ah, int nthreads);

/* bblas_dltrsm lower branches: trans 0 = N (L X = alpha B), 1 = T (L^T X = alpha B) */
void   lxo_dltrsm_lower(int trans, int bm, int bn, int m, int n, double alpha,
                        const double *A, int lda, double *B, int ldb, int nthreads);
void   lxo_sltrsm_lower(int trans, int bm, int bn, int m, int n, float alpha,
                        const float *A, int lda, float *B, int ldb, int nthreads);

/* dspdsolv_main: factor (RL) + forward + backward sweep, in place on A and B.
 * lda = ldb = m as in the reference.  Returns the factor's info. */
int    lxo_dspdsolv(int m, int n, int b, double *A, double *B, int nthreads);

/* Mixed-precision iterative refinement (FP32 tiled factor + solves, FP64
 * residual r = b - A x over the full stored A, FP64 update and test). */
typedef struct {
    int    max_sweeps;      /* refinement sweeps before falling back */
    double tol;             /* normwise backward error target */
    double stall_ratio;     /* ||r_k+1|| > stall_ratio ||r_k|| counts as a stall */
    int    nthreads;
} lxo_ir_opts;
typedef struct {
    int    sweeps;          /* corrections applied */
    int    fallback;        /* 1 when the FP64 factor produced the result */
    int    info;            /* factor info (0 = ok) */
    double berr;            /* final max over columns of ||r||/(||A|| ||x||), inf-norms */
} lxo_ir_status;
int    lxo_dposv_ir(int n, int nrhs, int b, const double *A, int lda,
                    const double *B, int ldb, double *X, int ldx,
                    const lxo_ir_opts *opts, lxo_ir_status *st);

/* Checks. */
/* norm1(tril(|A| - |L|)) / norm1(tril(|A|)) as in densutil.c:50-99 (lda = m there) */
double lxo_relerr_lower(int n, const double *A, int lda, const double *L, int ldl);
/* ||L L^T - A0||_inf / (||A0||_inf N eps)  (potrf_test.c:6-37) */
double lxo_potrf_resid(int n, const double *A0, int lda, const double *L, int ldl);
/* max_c ||B_c - A X_c||_inf / (||A||_inf ||X_c||_inf), full stored A */
double lxo_backward_error(int n, int nrhs, const double *A, int lda,
                          const double *X, int ldx, const double *B, int ldb);

double lxo_wtime(void);
int    lxo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
