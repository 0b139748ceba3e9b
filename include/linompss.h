/*
 * linompss.h -- C ABI of liblinompss.so, the MI355X (gfx950) implementation
 * of Linompss's tiled Cholesky / SPD solve / iterative refinement path.
 *
 * Matrices are column-major.  Only the lower triangle of A is read, and only
 * the lower triangle is overwritten with the factor L (A = L L^T); the strict
 * upper triangle is left as the caller passed it.
 *
 * Return values of the factor / solve entry points:
 *     0    success
 *    k>0   LAPACK-style info: the leading minor of order k is not positive
 *          definite (the reference discards dpotrf's info and returns 0,
 *          src/blocked_blas/tasks_potrf.c:22-23; no reference caller reads the
 *          value, so this is compatible)
 *    -i    argument i is invalid (LAPACK convention, 1-based)
 *    <= LINOMPSS_ERR_BASE  a device/runtime error (no GPU, HIP failure, the
 *          device scheduler gave up waiting, device memory exhausted)
 */
#ifndef LINOMPSS_H
#define LINOMPSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LINOMPSS_ERR_BASE   (-1000)
#define LINOMPSS_ERR_NODEV  (-1001)  /* no usable HIP device */
#define LINOMPSS_ERR_HIP    (-1002)  /* a HIP call failed */
#define LINOMPSS_ERR_NOMEM  (-1003)  /* device allocation failed */
#define LINOMPSS_ERR_SCHED  (-1004)  /* a work unit waited past the scheduler timeout */
#define LINOMPSS_ERR_PLAN   (-1005)  /* the planner rejected the shape */
#define LINOMPSS_ERR_IO     (-1006)  /* a matrix file could not be opened, written or parsed */
#define LINOMPSS_ERR_LIMIT  (-1007)  /* the size exceeds the fixed limit of a scratch-free entry point */

/* ---- Drop-in entry points (host memory, synchronous) -------------------- */

/* Replaces ompss_dchol_rl (src/cholesky/dense/ompss_dense_chol.h:38,
 * ompss_dense_chol.c:38-40 -> chol_rlmain.c:22-53).  Right-looking tiled
 * Cholesky with tile b; t is accepted and ignored, as in the reference. */
int ompss_dchol_rl(int n, int b, int t, double *A, int lda);
/* Replaces ompss_dchol_ll (ompss_dense_chol.h:34 -> chol_llmain.c:23-57). */
int ompss_dchol_ll(int n, int b, int t, double *A, int lda);
/* Replaces ompss_schol_rl (ompss_dense_chol.h:36; FP32 build of chol_rlmain.c). */
int ompss_schol_rl(int n, int b, int t, float *A, int lda);
/* Replaces ompss_schol_ll (ompss_dense_chol.h:32; FP32 build of chol_llmain.c). */
int ompss_schol_ll(int n, int b, int t, float *A, int lda);

/* Hypermatrix form: Ah is an mt x mt table of host tile pointers, tile (i, j)
 * at Ah[j * mt + i], each b x b column-major with leading dimension b; the
 * tiles i >= j are read and overwritten with the factor (the others may be
 * null).  Returns 0, the first non-positive global column (1-based), or a
 * negative error; t is accepted and ignored.
 * Replaces ompss_dchol_hrl (ompss_dense_chol.h:27, ompss_dense_chol.c:20-22 ->
 * chol_hrlmain.c:16-37). */
int ompss_dchol_hrl(int mt, int b, int t, double **Ah);
/* Replaces ompss_dchol_hll (ompss_dense_chol.h:19, ompss_dense_chol.c:12-14 ->
 * chol_hllmain.c:17-45). */
int ompss_dchol_hll(int mt, int b, int t, double **Ah);
/* Replaces ompss_schol_hrl (ompss_dense_chol.h:23; FP32 build of chol_hrlmain.c). */
int ompss_schol_hrl(int mt, int b, int t, float **Ah);
/* Replaces ompss_schol_hll (ompss_dense_chol.h:15; FP32 build of chol_hllmain.c). */
int ompss_schol_hll(int mt, int b, int t, float **Ah);
/* Replace the nested variants ompss_{d,s}chol_n{hrl,hll} (ompss_dense_chol.h:41-47,
 * ompss_dense_chol.c:48-62 -> chol_nrlmain.c:11, chol_nllmain.c:12): the same
 * factor of the same tiles (t only sub-blocks the reference's CPU tasks). */
int ompss_dchol_nhrl(int mt, int b, int t, double **Ah);
int ompss_dchol_nhll(int mt, int b, int t, double **Ah);
int ompss_schol_nhrl(int mt, int b, int t, float **Ah);
int ompss_schol_nhll(int mt, int b, int t, float **Ah);

/* ---- Standalone blocked BLAS-3 (host memory, synchronous) ---------------
 * The reference's flags are ompssblas_t (src/fp/fpmatr.h:14-17: a CBLAS enum
 * under MKL, a "N"/"T"-style string otherwise); here they are plain chars,
 * case-insensitive ('N'/'T', 'L'/'U', 'L'/'R' for side, 'N'/'U' for diag).
 * Block sizes are accepted (> 0) and ignored.  Return 0 or -i for argument i. */

/* C = alpha op(A) op(B) + beta C.  Replaces ompss_dgemm (src/gemm/ompss_mm.h:27,
 * ompss_mm.c:6-10 -> mm_main.c); linompss_sgemm replaces ompss_sgemm (:23). */
int linompss_dgemm(char transa, char transb, int b, int d, int c, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc);
int linompss_sgemm(char transa, char transb, int b, int d, int c, int m, int n, int k, float alpha,
                   const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc);
/* C = alpha op(A) op(A)^T + beta C on the uplo triangle (the other untouched).
 * Replaces ompss_dsyrk / ompss_ssyrk (src/syrk/ompss_syrk.h:22 / :20,
 * ompss_syrk.c:9-12 -> syrk_main.c). */
int linompss_dsyrk(char uplo, char trans, int b, int n, int k, double alpha, const double *A, int lda,
                   double beta, double *C, int ldc);
int linompss_ssyrk(char uplo, char trans, int b, int n, int k, float alpha, const float *A, int lda,
                   float beta, float *C, int ldc);
/* op(A) X = alpha B (side 'L', A m x m) or X op(A) = alpha B (side 'R', A n x n),
 * B m x n overwritten by X.  Replaces ompss_dtrsm / ompss_strsm
 * (src/trsm/ompss_trsm.h:23 / :20, ompss_trsm.c:4-12 -> trsm_main.c:6-17 ->
 * bblas_trsm.h:116-201); the reference leaves side 'R' unimplemented. */
int linompss_dtrsm(char side, char uplo, char trans, char diag, int m, int n, int b, double alpha,
                   const double *A, int lda, double *B, int ldb);
int linompss_strsm(char side, char uplo, char trans, char diag, int m, int n, int b, float alpha,
                   const float *A, int lda, float *B, int ldb);

/* ---- The reference's own BLAS-3 and tile-task symbols --------------------
 * Flags as the reference passes them: 8-byte unions by value
 * (src/fp/fpmatr.h:14-35), holding a CBLAS enumeration (its MKL build,
 * fpmatr.c:14-27: 111 NoTrans, 112 Trans, 121 Upper, 122 Lower, 131 NonUnit,
 * 132 Unit, 141 Left, 142 Right) or pointing at "N"/"T"/"L"/"U"/"R" (its
 * Fortran-BLAS build, fpmatr.c:36-47); both are accepted.  The descriptor
 * instances below are the Fortran-BLAS flavour.                            */
#ifndef __FPMATR_H__
typedef union {
    int e;
    char *s;
} ompssblas_t;
typedef union {
    char e;
    char *s;
} ompsslapack_t;
#endif
extern ompssblas_t OMPSSBLAS_ROWMAJOR, OMPSSBLAS_COLMAJOR, OMPSSBLAS_TRANSP, OMPSSBLAS_NTRANSP,
    OMPSSBLAS_UPPERTRIANG, OMPSSBLAS_LOWERTRIANG, OMPSSBLAS_NTRIANG, OMPSSBLAS_DIAGUNIT, OMPSSBLAS_NDIAGUNIT,
    OMPSSBLAS_LEFT, OMPSSBLAS_RIGHT;
extern ompsslapack_t OMPSSLAPACK_UPPERTRIANG, OMPSSLAPACK_LOWERTRIANG;

/* Replace ompss_dgemm / ompss_sgemm (src/gemm/ompss_mm.h:27 / :23): the
 * linompss_{d,s}gemm above with the reference's descriptors; returns 0, -1 /
 * -2 for an unreadable flag, or -i as linompss_dgemm does. */
int ompss_dgemm(ompssblas_t transa, ompssblas_t transb, int b, int d, int c, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc);
int ompss_sgemm(ompssblas_t transa, ompssblas_t transb, int b, int d, int c, int m, int n, int k, float alpha,
                const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc);
/* Replace ompss_dsyrk / ompss_ssyrk (src/syrk/ompss_syrk.h:22 / :20) and
 * ompss_dtrsm / ompss_strsm (src/trsm/ompss_trsm.h:23 / :20).  void as in the
 * reference: an argument or device error is printed to stderr. */
void ompss_dsyrk(ompssblas_t uplo, ompssblas_t trans, int b, int n, int k, double alpha, double *A, int lda,
                 double beta, double *C, int ldc);
void ompss_ssyrk(ompssblas_t uplo, ompssblas_t trans, int b, int n, int k, float alpha, float *A, int lda,
                 float beta, float *C, int ldc);
void ompss_dtrsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int m, int n, int b,
                 double alpha, double *A, int lda, double *B, int ldb);
void ompss_strsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int m, int n, int b,
                 float alpha, float *A, int lda, float *B, int ldb);
/* The tile tasks of libbblas (src/blocked_blas/tasks_potrf.h:23, tasks_trsm.h:24,
 * tasks_syrk.h:21, task_gemm.h:27, and the s twins): one tile operation on
 * host memory, run on the device and completed before the call returns (the
 * reference spawns an OmpSs task; p, its priority, is ignored).  task_dpotrf
 * factors the uplo triangle in place and, as the reference does, discards
 * LAPACK's info. */
void task_dpotrf(ompsslapack_t uplo, int n, double *A, int lda, int p);
void task_spotrf(ompsslapack_t uplo, int n, float *A, int lda, int p);
void task_dtrsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int bm, int bn, double alpha,
                double *A, int lda, double *B, int ldb, int p);
void task_strsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int bm, int bn, float alpha,
                float *A, int lda, float *B, int ldb, int p);
void task_dsyrk(ompssblas_t uplo, ompssblas_t trans, int n, int k, double alpha, double *A, int lda, double beta,
                double *C, int ldc, int p);
void task_ssyrk(ompssblas_t uplo, ompssblas_t trans, int n, int k, float alpha, float *A, int lda, float beta,
                float *C, int ldc, int p);
void task_dgemm(ompssblas_t transa, ompssblas_t transb, int bm, int bn, int bk, double alpha, double *A, int lda,
                double *B, int ldb, double beta, double *C, int ldc, int p);
void task_sgemm(ompssblas_t transa, ompssblas_t transb, int bm, int bn, int bk, float alpha, float *A, int lda,
                float *B, int ldb, float beta, float *C, int ldc, int p);

/* ---- Dense matrices in the Octave text format (A.dat / RHS.dat of the
 * refinement app, src/itref/itref_setup.c:98-121, 155-199) ------------------
 * Writer: replaces fprint_dense2mm / print_dense2mm (src/matutil/matfprint.c:
 * 94-124): "# name:", "# type: matrix", "# rows:", "# columns:" header, then
 * the m x n values row-major, one "%.16e \n" per line.  Returns 0, -i for
 * argument i, or LINOMPSS_ERR_IO. */
int linompss_fprint_ddense2mm(const char *fname, const char *name, int m, int n, const double *A, int lda);
int linompss_fprint_sdense2mm(const char *fname, const char *name, int m, int n, const float *A, int lda);
/* Raw reader: replaces read_mm2ddense / read_mm2sdense (src/matutil/matfread.c:
 * 13-27): up to count values in file order, '#' lines skipped, a line that
 * does not parse repeating the last value.  Returns the number of values
 * stored, -i for argument i, or LINOMPSS_ERR_IO. */
int64_t linompss_read_mm2ddense(const char *fname, int64_t count, double *out);
int64_t linompss_read_mm2sdense(const char *fname, int64_t count, float *out);
/* The header's rows and columns; then the values into a column-major m x n
 * array (ld lda) in the layout the writer used (the reference's setup reads
 * the row-major values straight into column-major storage).  No counterpart
 * of its own in the reference. */
int linompss_dense2mm_dims(const char *fname, int *m, int *n);
int linompss_read_ddense2mm(const char *fname, int m, int n, double *A, int lda);
int linompss_read_sdense2mm(const char *fname, int m, int n, float *A, int lda);

/* Replaces dspdsolv_main (src/spdsolv/spdsolv_main.h:20, spdsolv_main.c:10-20):
 * RL factor of the m x m matrix A (lda = m), then the forward and backward
 * sweeps on the m x n right-hand sides B (ldb = m), in place. */
int dspdsolv_main(int m, int n, int b, double *A, double *B);
/* Replaces sspdsolv_main (spdsolv_main.h:19). */
int sspdsolv_main(int m, int n, int b, float *A, float *B);

/* Replaces the lower branches of bblas_dltrsm (src/blocked_blas/bblas_trsm.h
 * :116-160), a header-inline routine in the reference: B := alpha op(L)^{-1} B
 * with op = N ('N') or T ('T'), L the m x m lower factor in A, B m x n.
 * uplo must be 'L' and diag 'N' (the only branches on the SPD path). */
int linompss_dltrsm(char uplo, char trans, char diag, int bm, int bn, int m, int n, double alpha,
                    const double *A, int lda, double *B, int ldb);
int linompss_sltrsm(char uplo, char trans, char diag, int bm, int bn, int m, int n, float alpha,
                    const float *A, int lda, float *B, int ldb);

/* Mixed-precision iterative refinement.  The reference's itref has no C API
 * (only a CLI, src/itref/itref_config.c:130-223); this entry point keeps its
 * loop contract (itref.c:999-1101: x0 = solve(b); r = b - A x in FP64;
 * repeat { test; d = solve(r); x += d; r = b - A x }) with the inner solver
 * an FP32 tiled Cholesky solve.  The normwise backward error
 * max_c ||r_c||_inf / (||A||_inf ||x_c||_inf) is tested before each
 * correction; when it stalls or the sweep cap is reached the FP64 factor
 * solves the system instead. A (lower triangle, symmetric) is not modified. */
typedef struct {
    int max_sweeps;       /* corrections before falling back (default 30) */
    double tol;           /* backward-error target (default 1e-14) */
    double stall_ratio;   /* ||r_k+1|| > stall_ratio ||r_k|| is a stall (default 0.5) */
    int update_prec;      /* FP32 factor's trailing updates: 0 FP32 (default), 1 split BF16,
                             2 split FP16 (see linompss_exec_opts.update_prec); with 2 the
                             FP32 copy of A is equilibrated first, D A D with D = diag of
                             powers of two making every diagonal entry lie in [1/4, 1), and
                             each correction solves with it: d = D (D A D)^-1 D r */
} linompss_ir_opts;

typedef struct {
    int sweeps;           /* corrections applied */
    int fallback;         /* 1 when the FP64 factor produced the result */
    int info;             /* info of the factor that produced the result */
    double berr;          /* final normwise backward error */
    double t_factor_s;    /* device time of the low-precision factor */
    double t_total_s;     /* wall time of the whole solve */
} linompss_ir_status;

int linompss_dposv_ir(int n, int nrhs, int b, const double *A, int lda, const double *B, int ldb,
                      double *X, int ldx, const linompss_ir_opts *opts, linompss_ir_status *st);

/* ---- Device-memory entry points ------------------------------------------
 * Pointers are HIP device pointers; stream is a hipStream_t (NULL = the null
 * stream).  Work is enqueued on the stream and the call returns after it
 * completed (the info value needs the result).  dtype: 0 = FP64, 1 = FP32.
 * variant: 0 = right-looking (CHOL_RL), 1 = left-looking (CHOL_LL).        */

typedef struct {
    int policy;           /* unit order: 0 reference priorities, 1 critical path (default) */
    int unit_m, unit_n;   /* GEMM/SYRK work-unit tile (default 128 x 128) */
    int wg_per_cu;        /* resident workgroups per CU (0 = occupancy limit) */
    double timeout_s;     /* per-dependency wait limit on the device (default 30 s) */
    int fused_solve;      /* posv: 1 = factor and sweeps in one persistent launch (the full
                             task DAG); 0 (default) = factor launch + sweep kernels */
    int update_prec;      /* FP32 factor's trailing SYRK/GEMM: 0 FP32 matrix cores (default);
                             1 BF16 matrix cores on split operands (each FP32 input as
                             hi + mid + lo BF16, six products, FP32 accumulation: FP32-level
                             accuracy at the BF16 rate); 2 FP16 matrix cores on split
                             operands (x * 2^12 as hi + lo FP16, three products, half the
                             matrix-core work of 1 at about 4x its rounding error per
                             product; needs every entry of L within 16 in magnitude, which
                             a diagonal of at most 256 guarantees: past it the factor
                             reports a non-positive pivot); ignored for FP64 */
} linompss_exec_opts;

typedef struct {
    double exec_ms;       /* device time of the last executor launch (HIP events) */
    int units;            /* work units in that launch */
    int tasks;            /* reference-level tile tasks in that launch */
    int grid;             /* persistent workgroups launched */
    double sim_ms;        /* planner's simulated makespan */
} linompss_exec_stats;

int linompss_chol_dev(int dtype, int variant, int n, int b, void *dA, int64_t lda, void *stream,
                      const linompss_exec_opts *opts);
/* The factor of a hypermatrix in device memory: dTiles is a DEVICE table of
 * mt * mt device tile pointers, tile (i, j) at dTiles[j * mt + i], b x b with
 * leading dimension b (tiles i >= j read and written).  Same plans and units
 * as linompss_chol_dev on the n = mt * b matrix, each unit addressed inside
 * its tile.  Device form of ompss_{d,s}chol_{hrl,hll} (ompss_dense_chol.h:15-27;
 * chol_hrlmain.c:16-37, chol_hllmain.c:17-45). */
int linompss_chol_tiles_dev(int dtype, int variant, int mt, int b, void *const *dTiles, void *stream,
                            const linompss_exec_opts *opts);
/* Device forms of the BLAS-3 entry points above (dtype 0 FP64, 1 FP32; alpha
 * and beta converted to the dtype): the same flags and argument checks. */
int linompss_gemm_dev(int dtype, char transa, char transb, int m, int n, int k, double alpha, const void *dA,
                      int64_t lda, const void *dB, int64_t ldb, double beta, void *dC, int64_t ldc, void *stream);
int linompss_syrk_dev(int dtype, char uplo, char trans, int n, int k, double alpha, const void *dA, int64_t lda,
                      double beta, void *dC, int64_t ldc, void *stream);
int linompss_trsm_blas_dev(int dtype, char side, char uplo, char trans, char diag, int m, int n, double alpha,
                           const void *dA, int64_t lda, void *dB, int64_t ldb, void *stream);
int linompss_posv_dev(int dtype, int n, int nrhs, int b, void *dA, int64_t lda, void *dX, int64_t ldx,
                      void *stream, const linompss_exec_opts *opts);
int linompss_potrs_dev(int dtype, int n, int nrhs, int b, const void *dL, int64_t lda, void *dX,
                       int64_t ldx, void *stream, const linompss_exec_opts *opts);
/* linompss_potrs_dev that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  The same sweep kernels run
 * on a workspace in a fixed device array of the library (under 24 MiB per
 * device, held until the process ends), so n <= 32768 (LINOMPSS_ERR_LIMIT
 * beyond), any nrhs, opts->fused_solve must be 0 (-10 otherwise).  Every call
 * on a device shares that array: calls that overlap on one device race (the
 * library serialises its own calls).  Results equal linompss_potrs_dev's. */
int linompss_potrs_dev_noalloc(int dtype, int n, int nrhs, int b, const void *dL, int64_t lda, void *dX,
                               int64_t ldx, void *stream, const linompss_exec_opts *opts);
/* One triangular sweep on device memory: X := op(L)^-1 X, op = N (trans 0)
 * or T (trans 1), L lower (the device twin of linompss_dltrsm). */
int linompss_trsm_dev(int dtype, int trans, int n, int nrhs, const void *dL, int64_t ldl, void *dX, int64_t ldx,
                      void *stream);
int linompss_dposv_ir_dev(int n, int nrhs, int b, const double *dA, int64_t lda, const double *dB,
                          int64_t ldb, double *dX, int64_t ldx, void *stream, const linompss_ir_opts *opts,
                          linompss_ir_status *st);

/* ---- Batched systems ------------------------------------------------------
 * Many independent SPD systems of uneven order (0 <= n_i <= 512, one
 * right-hand side each), each solved as linompss_dposv_ir_dev solves one
 * system (FP32 Cholesky factor, FP64 residual / update / backward-error test,
 * FP64-factor fallback on a stall; itref.c:999-1101), in ONE persistent
 * launch: a workgroup takes a whole system at a time from a ticket queue.  The
 * reference has no batched entry point (its dspdsolv_main, spdsolv_main.c:10-20,
 * solves one system per call); these calls stand where a caller would loop
 * over dspdsolv_main.
 *
 * A plan fixes the batch (orders, device pointers, leading dimensions), picks
 * the schedule (workgroups per CU, system order) -- measured on the batch
 * itself when left to the autotuner, which runs the batch and so writes X --
 * and captures the launch into a hipGraph.  linompss_batch_run replays it on
 * a stream and waits; A and B are not modified.  Runs of one plan must not
 * overlap (they share its queue counter and workspace).                     */
typedef struct {
    int max_sweeps;       /* corrections before falling back (default 30) */
    double tol;           /* backward-error target (default 1e-14) */
    double stall_ratio;   /* ||r_k+1|| > stall_ratio ||r_k|| is a stall (default 0.5) */
    int wg_per_cu;        /* resident workgroups per CU: 1 or 2; 0 = autotuned */
    int order;            /* 0 largest system first, 1 as given; -1 = autotuned */
    int use_graph;        /* 1: runs replay a captured hipGraph; 0: plain launches */
    int panel;            /* FP32 factor panel width: 64 or 128 for every system; 0 =
                             per size class (orders 1-64, 65-128, ...): measured on the
                             batch's own systems when the plan tunes its schedule, else
                             128 above order 256 and 64 below */
} linompss_batch_opts;

typedef struct {
    int sweeps;           /* corrections applied */
    int fallback;         /* 1 when the FP64 factor produced the result */
    int info;             /* info of the factor that produced the result */
    int reserved;
    double berr;          /* final normwise backward error */
} linompss_batch_status;

typedef struct linompss_batch_plan linompss_batch_plan;

/* n, dA, lda, dB, dX: host arrays of count entries (the pointers are device
 * pointers).  opts NULL = {30, 1e-14, 0.5, 0, -1, 1, 0}.  Returns 0, -i for an
 * invalid argument i, or LINOMPSS_ERR_*. */
int linompss_batch_plan_create(int count, const int *n, const double *const *dA, const int64_t *lda,
                               const double *const *dB, double *const *dX, const linompss_batch_opts *opts,
                               linompss_batch_plan **plan);
/* st: count entries, or NULL. */
int linompss_batch_run(linompss_batch_plan *plan, void *stream, linompss_batch_status *st);
int linompss_batch_plan_info(const linompss_batch_plan *plan, int *grid, int *wg_per_cu, int *order,
                             double *tune_ms);
/* Bit c of *mask: systems of order 64c+1 .. 64c+64 factor in 128-column
 * panels (else 64).  No counterpart in the reference. */
int linompss_batch_plan_panels(const linompss_batch_plan *plan, unsigned *mask);
int linompss_batch_plan_destroy(linompss_batch_plan *plan);
/* One-shot: plan (opts NULL = one workgroup per CU, largest first, no graph),
 * run, destroy. */
int linompss_dposv_ir_batched_dev(int count, const int *n, const double *const *dA, const int64_t *lda,
                                  const double *const *dB, double *const *dX, void *stream,
                                  const linompss_batch_opts *opts, linompss_batch_status *st);

/* ---- Adjoints (reverse mode) ----------------------------------------------
 * No counterpart in the reference (its factor is task_dpotrf,
 * tasks_potrf.c:20-24, and its solves bblas_dltrsm, bblas_trsm.h:116-160,
 * with no derivatives); these give the gradients a differentiable caller
 * needs, in FP64.
 *
 * Cholesky adjoint: given L (A = L L^T, lower) and Lbar (the adjoint of L,
 * lower), overwrites Lbar with Abar in the lower representation
 * (dLoss = sum_{i>=j} Abar_ij dA_ij; the symmetric gradient is Abar with its
 * off-diagonal entries halved and mirrored).  Blocked reverse mode over
 * 64-column blocks (I. Murray, "Differentiation of the Cholesky
 * decomposition", 2016), ~2n^3/3 flops on the FP64 matrix cores.
 * logdet_weight w adds the adjoint of w log det A (Lbar_ii += 2 w / L_ii)
 * before the sweep, the term every likelihood carries.
 * linompss_dpotrf_adjoint takes A instead of L and recomputes the factor on
 * the device (tile b) rather than keeping it from the forward pass.       */
int linompss_dpotrf_adjoint(int n, int b, const double *A, int lda, double *Lbar, int ldlbar, double logdet_weight);
int linompss_dpotrf_adjoint_dev(int n, const double *dL, int64_t ldl, double *dLbar, int64_t ldlbar,
                                double logdet_weight, void *stream);

/* Triangular-solve adjoint: X = op(L)^-1 B (op = 'N' or 'T', L lower n x n,
 * nrhs columns).  Xbar (in: the adjoint of X) is overwritten with
 * Bbar = op(L)^-T Xbar, and Lbar (when not NULL) is updated with
 * -tril(Bbar X^T) (op = N) or -tril(X Bbar^T) (op = T); the device entry
 * point overwrites Lbar's lower triangle with that term when accumulate is 0. */
int linompss_dtrsm_adjoint(char trans, int n, int nrhs, const double *L, int ldl, const double *X, int ldx,
                           double *Xbar, int ldxbar, double *Lbar, int ldlbar);
int linompss_dtrsm_adjoint_dev(int trans, int n, int nrhs, const double *dL, int64_t ldl, const double *dX,
                               int64_t ldx, double *dXbar, int64_t ldxbar, double *dLbar, int64_t ldlbar,
                               int accumulate, void *stream);

/* log det A = 2 sum_i log L_ii from the device factor L (host result). */
int linompss_dlogdet_dev(int n, const double *dL, int64_t ldl, double *logdet, void *stream);

/* Normwise backward error max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf),
 * computed in FP64 on the device (A symmetric, lower triangle read; any nrhs).
 * A column with a zero residual counts as backward error 0 (b_c = 0 gives
 * x_c = 0).  Returns 0 or -i for argument i. */
int linompss_dberr_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dX, int64_t ldx,
                       const double *dB, int64_t ldb, double *berr);

/* ---- Factor autotuner (no counterpart in the reference, whose tile size
 * and task shapes are fixed by its caller and its code).  A configuration is
 * the tile size plus the shapes of the work units the planner cuts the tile
 * tasks into (chol_rlmain.c:25-50 / chol_llmain.c:26-53 stay the task graph):
 *   potrf_sub       diagonal tiles wider than this factor as a sub-DAG of
 *                   potrf_sub-wide tiles (multiple of 64)
 *   gemm_wide_m     rows of a trailing-update unit off the lookahead column
 *                   (multiple of 128)
 *   gemm_wide_n     columns of a trailing-update unit off the lookahead column
 *                   (multiple of 128)
 *   trsm_rows_wide  rows of a panel-solve unit off the lookahead tile
 *                   (multiple of 64)
 *   wide_tail       the last wide_tail panels use 128-column update units
 *                   (-1: half of the panels)
 *   update_group    g (1 to 8): right-looking plans take the steps in groups
 *                   of g and give the tiles right of a group's columns one
 *                   pass for the group's g panels (each such tile read and
 *                   written once per g GEMM / SYRK tasks; the task graph is
 *                   unchanged); 1: one pass per task.  Hypermatrix plans
 *                   always use 1.
 * linompss_tune_factor times the factor of an n x n diagonally dominant SPD
 * matrix it generates on the device (unit diagonal, off-diagonal entries in
 * [0, 1/n)) under candidate configurations, by coordinate descent from the
 * library's defaults (each candidate the fastest of reps launches, reps <= 0
 * means 3), and records the fastest for (dtype, variant, n, best->b,
 * update_prec): every later factor of that shape (linompss_chol_dev,
 * linompss_posv_dev, the refinement driver's low-precision factor, the
 * ompss_* entry points) uses it.  b > 0 keeps the tile size; b = 0 searches
 * it among 384, 512, 768 and 1024, and 128 and 256 when n <= 16384 (those
 * <= n).  Needs n*n elements of
 * temporary device memory.  update_prec as in linompss_exec_opts (FP32 only).
 * tried (may be NULL) receives up to max_tried configurations in the order
 * they were timed, *ntried their count.  Returns 0, -i for argument i, or a
 * LINOMPSS_ERR_* code.
 * linompss_set_factor_config records a configuration (e.g. one saved from an
 * earlier run) with the same effect; linompss_get_factor_config returns the
 * recorded one (1 if there is one, 0 if the shape uses the defaults, which
 * are then written to *cfg with ms = 0); linompss_clear_factor_configs forgets
 * them all.  LINOMPSS_* environment knobs, where set, take precedence. */
typedef struct {
    int b;
    int potrf_sub;
    int gemm_wide_m;
    int gemm_wide_n;
    int trsm_rows_wide;
    int wide_tail;
    int update_group;
    double ms;  /* factor launch time measured for this configuration */
} linompss_factor_config;

int linompss_tune_factor(int dtype, int variant, int n, int b, int update_prec, int reps,
                         linompss_factor_config *best, linompss_factor_config *tried, int max_tried, int *ntried);
int linompss_set_factor_config(int dtype, int variant, int n, int update_prec, const linompss_factor_config *cfg);
int linompss_get_factor_config(int dtype, int variant, int n, int b, int update_prec, linompss_factor_config *cfg);
int linompss_clear_factor_configs(void);

/* ---- Gaussian-process kernel on the device (no counterpart in the
 * reference; it feeds the factor and the adjoints above so that a GP
 * likelihood's forward and backward passes keep K on the device).
 * linompss_gp_kernel_dev: K_ij = sf2 exp(-|x_i - x_j|^2 / (2 ell^2)) +
 * noise delta_ij for the n points X (n x d column-major, ldx >= n, d <= 16),
 * both triangles of K (ldk >= n) written.
 * linompss_gp_grad_dev: with Kbar the adjoint of K in the lower representation
 * (dLoss = sum_{i>=j} Kbar_ij dK_ij, as linompss_dpotrf_adjoint_dev writes
 * it), grads[0] = dLoss/d ell, grads[1] = dLoss/d sf2, grads[2] = dLoss/d noise
 * (host array of 3), and, when dXbar is not NULL, dXbar = dLoss/dX (n x d,
 * device).  Sums in a fixed order (deterministic).  Both return 0, -i for
 * argument i, or a LINOMPSS_ERR_* code, after the stream has completed. */
int linompss_gp_kernel_dev(int n, int d, const double *dX, int64_t ldx, double ell, double sf2, double noise,
                           double *dK, int64_t ldk, void *stream);
int linompss_gp_grad_dev(int n, int d, const double *dX, int64_t ldx, double ell, double sf2, const double *dKbar,
                         int64_t ldkb, double *grads, double *dXbar, int64_t ldxb, void *stream);

/* Statistics of the last executor launch on this thread. */
int linompss_last_exec_stats(linompss_exec_stats *st);

/* Task graph of a plan (for checking against the reference's recorded DAG,
 * visualize/TDG/dpotrf_tdg.dot).  what: 0 RL factor, 1 LL factor, 2 RL factor
 * + both sweeps.  types[t] in {0 POTRF, 1 TRSM, 2 GEMM, 3 SYRK, 4 sweep TRSM,
 * 5 sweep GEMM}; coords[3t..3t+2] = tile (i, j, k); edges[2e], edges[2e+1] =
 * (pred, succ).  Returns 0, or -1 when a buffer is too small (counts are still
 * stored). */
int linompss_plan_tdg(int what, int n, int b, int nrhs, int *types, int *coords, int max_tasks, int *edges,
                      int max_edges, int *ntasks, int *nedges);

/* Unit count and simulated makespan of a plan (no device needed). */
int linompss_plan_info(int what, int n, int b, int nrhs, const linompss_exec_opts *opts, int *units,
                       int *tasks, double *sim_ms);

const char *linompss_version(void);

/* ---- Workload generators (device memory; used by the benchmark and tests).
 * linompss_gen_gauss_spd: A = G G^T / n + alpha I (both triangles stored),
 * G i.i.d. N(0,1) from Philox4x32-10 keyed by seed, alpha = lambda_max /
 * (kappa - 1) with lambda_max(G G^T / n) from Lanczos (largest Ritz value), so that
 * cond2(A) ~ kappa.  Needs n*n*8 bytes of temporary device memory.
 * linompss_gen_uniform: X (m x n, dtype 0 FP64 / 1 FP32) uniform in [lo, hi).
 * linompss_gen_gauss: G (m x n, FP64) i.i.d. N(0,1), element e = j*m + i
 * taken from Philox block e/2 (Box-Muller pair).                             */
int linompss_gen_gauss_spd(int n, uint32_t seed, double kappa, double *dA, int64_t lda, double *alpha,
                           double *lambda_max);
int linompss_gen_uniform(int dtype, int m, int n, uint32_t seed, double lo, double hi, void *dX, int64_t ldx);
int linompss_gen_gauss(int m, int n, uint32_t seed, double *dG, int64_t ldg);
/* The refinement app's default "COVAR" matrix: GENMAT_COVARIANT
 * (src/matutil/genmat.c:263-302, reached through dgenmat_config,
 * genmat_config.c:34-133, with p = k = 2 by default): a_ij = |i - j|^-p off
 * the diagonal, a_jj = 1 + (j + 1)^(1/k), both triangles, n x n with leading
 * dimension lda, in device memory.  Returns 0, -i for argument i (p <= 0 is
 * refused, as the reference returns NULL), or LINOMPSS_ERR_HIP. */
int linompss_gen_covariant(int n, double p, double k, double *dA, int64_t lda);
/* The reference's own drand48 generators, bit for bit, in device memory
 * (dtype 0 FP64, 1 FP32; seed as srand48 takes it, the reference passes
 * time(NULL)): GENMAT_SPD (src/matutil/genmat.c:49-70), a_ij = a_ji drawn
 * column by column over i >= j with n added to the diagonal, and GENMAT
 * (genmat.c:106-128), X(i, j) the (j m + i + 1)-th draw.  Return 0, -i for
 * argument i, or LINOMPSS_ERR_HIP. */
int linompss_gen_drand48_spd(int dtype, int n, long seed, void *dA, int64_t lda);
int linompss_gen_drand48(int dtype, int m, int n, long seed, void *dX, int64_t ldx);
/* The factor check of the dchol CLI (chol_check.c:12-42): L2 = L L^T from the
 * lower triangle of dL, then DMAT_RELERR (densutil.c:50-99),
 * max_j sum_{i>=j} | |L2_ij| - |A_ij| | / max_j sum_{i>=j} |L2_ij|, in
 * *relerr (column sums in FP64).  Allocates 2 n^2 elements of scratch. */
int linompss_relerr_lower_dev(int dtype, int n, const void *dL, int64_t ldl, const void *dA, int64_t lda,
                              double *relerr);

/* ---- Device-memory plumbing (so C and ctypes callers need no other HIP
 * binding).  kind: 1 host->device, 2 device->host, 3 device->device.       */
int linompss_device_count(void);
void *linompss_dev_alloc(size_t bytes);
int linompss_dev_free(void *p);
int linompss_memcpy(void *dst, const void *src, size_t bytes, int kind);
int linompss_dev_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* LINOMPSS_H */
