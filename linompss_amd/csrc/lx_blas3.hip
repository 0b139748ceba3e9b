// lx_blas3.hip -- the reference's standalone blocked BLAS-3 entry points on
// the GPU:
//   GEMM  C = alpha op(A) op(B) + beta C   src/gemm/ompss_mm.h:23-27 (ompss_{s,d}gemm)
//   SYRK  C = alpha op(A) op(A)^T + beta C, one triangle
//                                           src/syrk/ompss_syrk.c:9-12 -> syrk_main.c
//   TRSM  op(A) X = alpha B  (left) or X op(A) = alpha B (right), A triangular
//                                           src/trsm/ompss_trsm.c:9-12 -> trsm_main.c:6-17
//                                           -> bblas_trsm.h:116-201
// The reference cuts each call into tile tasks with its block sizes (b, d, c);
// here GEMM and SYRK are one launch of the 64 x 64-tile matrix-core kernel
// (lx_gemm64.hpp) and the block sizes are accepted and ignored, as t is in
// the Cholesky entry points.  TRSM runs on the triangular sweeps: the
// operator is first written as a lower triangle Lt (the transpose of an upper
// A, ones on the diagonal for a unit A) so that every case is a forward
// (Lt Y = B) or backward (Lt^T Y = B) sweep; a right-side solve transposes B
// and solves on the left.  The reference's TRSM has no right side
// (trsm_main.c:13-15 prints "not yet implemented" and returns); here it is
// solved.  With more than 16 right-hand sides (one sweep pass) the solve is
// blocked over 64-row steps on the matrix cores instead (lx_b3_trsm_blocked).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>

#include "linompss.h"
#include "lx_gemm64.hpp"
#include "lx_internal.hpp"

// Lt (n x n, ld n) := the lower-triangular operator: tril(A) for a lower A,
// (triu(A))^T for an upper A; ones on the diagonal when unit; zeros above.
template <typename T>
__global__ void lx_b3_lower_kernel(int n, const T *__restrict__ A, int64_t lda, int upper, int unit,
                                   T *__restrict__ Lt)
{
    const int64_t tot = (int64_t)n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % n), c = (int)(e / n);
        T v = T(0);
        if (r == c) v = unit ? T(1) : A[(int64_t)c * lda + r];
        else if (r > c) v = upper ? A[(int64_t)r * lda + c] : A[(int64_t)c * lda + r];
        Lt[e] = v;
    }
}

// B (m x n) *= alpha (alpha == 0 writes zeros without reading B).
template <typename T>
__global__ void lx_b3_scale_kernel(int m, int n, T alpha, T *__restrict__ B, int64_t ldb)
{
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        T *p = B + (e / m) * ldb + e % m;
        *p = alpha == T(0) ? T(0) : alpha * *p;
    }
}

// D (cols x rows, ld ldd) := S^T (S rows x cols, ld lds), through 32 x 32 LDS
// tiles so that both the reads and the writes run down columns.
template <typename T>
__global__ void __launch_bounds__(256) lx_b3_transpose_kernel(int rows, int cols, const T *__restrict__ S, int64_t lds,
                                                              T *__restrict__ D, int64_t ldd)
{
    __shared__ T t[32][33];
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
        const int r = r0 + tx, c = c0 + ty + q;
        t[ty + q][tx] = (r < rows && c < cols) ? S[(int64_t)c * lds + r] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
        const int c = c0 + tx, r = r0 + ty + q;  // D[c][r] = S[r][c]
        if (r < rows && c < cols) D[(int64_t)r * ldd + c] = t[tx][ty + q];
    }
}

// Z + 4096 J := the inverse of the 64 x 64 diagonal block J of the lower
// triangle Lt (n x n, ld ld); identity past the matrix's order.  One workgroup
// per block.
template <typename T>
__global__ void __launch_bounds__(LX_WG_THREADS) lx_b3_inv_kernel(int n, const T *__restrict__ Lt, int64_t ld,
                                                                  T *__restrict__ Z)
{
    __shared__ T sL[64][65], sX[64][65];
    const int tid = lx_tid(), j0 = 64 * (int)blockIdx.x, jb = min(64, n - j0);
    for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
        const int r = e & 63, c = e >> 6;
        sL[r][c] = (r < jb && c < jb && r >= c) ? Lt[(int64_t)(j0 + c) * ld + j0 + r] : (r == c ? T(1) : T(0));
        sX[r][c] = T(0);
    }
    __syncthreads();
    lx_tri_inv64<T>(sL, sX);  // ends with a barrier
    for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
        const int r = e & 63, c = e >> 6;
        Z[(int64_t)blockIdx.x * 4096 + c * 64 + r] = sX[r][c];
    }
}

static int lx_b3_cus()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
    }
    return cus;
}

static inline int lx_b3_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }
static inline int lx_b3_flag(char f, const char *ok)  // index of f in ok (case-insensitive), -1 if absent
{
    for (int i = 0; ok[i]; ++i)
        if (f == ok[i] || f == ok[i] + ('a' - 'A')) return i;
    return -1;
}

// Split-K partials: up to 16 M x N slices (scratch slot 11, grow-only).
template <typename T>
static int lx_b3_gemm_dev(int ta, int tb, int m, int n, int k, T alpha, const T *A, int64_t lda, const T *B,
                          int64_t ldb, T beta, T *C, int64_t ldc, int uplo, hipStream_t s)
{
    size_t pe = 0;
    T *part = nullptr;
    const int64_t tiles = (int64_t)((m + 63) / 64) * ((n + 63) / 64);
    if (tiles < 2 * lx_b3_cus() && k >= 512) {
        pe = std::min<size_t>((size_t)16 * m * n, (size_t)64 << 20);
        part = (T *)lx_scratch(11, sizeof(T) * pe);
        if (!part) pe = 0;
    }
    if (lx_gemm64<T>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, uplo, part, pe, lx_b3_cus(), s))
        return LINOMPSS_ERR_HIP;
    LX_HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

extern "C" int linompss_gemm_dev(int dtype, char transa, char transb, int m, int n, int k, double alpha,
                                 const void *dA, int64_t lda, const void *dB, int64_t ldb, double beta, void *dC,
                                 int64_t ldc, void *stream)
{
    const int ta = lx_b3_flag(transa, "NT"), tb = lx_b3_flag(transb, "NT");
    if (dtype != 0 && dtype != 1) return -1;
    if (ta < 0) return -2;
    if (tb < 0) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    if (lda < std::max(1, ta ? k : m)) return -9;
    if (ldb < std::max(1, tb ? n : k)) return -11;
    if (ldc < std::max(1, m)) return -14;
    if (m == 0 || n == 0) return 0;
    std::lock_guard<std::mutex> g(lx_dev_mutex());
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
        return lx_b3_gemm_dev<double>(ta, tb, m, n, k, alpha, (const double *)dA, lda, (const double *)dB, ldb, beta,
                                      (double *)dC, ldc, 0, s);
    return lx_b3_gemm_dev<float>(ta, tb, m, n, k, (float)alpha, (const float *)dA, lda, (const float *)dB, ldb,
                                 (float)beta, (float *)dC, ldc, 0, s);
}

extern "C" int linompss_syrk_dev(int dtype, char uplo, char trans, int n, int k, double alpha, const void *dA,
                                 int64_t lda, double beta, void *dC, int64_t ldc, void *stream)
{
    const int up = lx_b3_flag(uplo, "LU"), tr = lx_b3_flag(trans, "NT");
    if (dtype != 0 && dtype != 1) return -1;
    if (up < 0) return -2;
    if (tr < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    if (lda < std::max(1, tr ? k : n)) return -8;
    if (ldc < std::max(1, n)) return -11;
    if (n == 0) return 0;
    std::lock_guard<std::mutex> g(lx_dev_mutex());
    hipStream_t s = (hipStream_t)stream;
    // op(A) op(A)^T: N -> A (n x k) times A^T; T -> A^T times A (A k x n)
    const int ta = tr ? 1 : 0, tb = tr ? 0 : 1;
    if (dtype == 0)
        return lx_b3_gemm_dev<double>(ta, tb, n, n, k, alpha, (const double *)dA, lda, (const double *)dA, lda, beta,
                                      (double *)dC, ldc, up ? 2 : 1, s);
    return lx_b3_gemm_dev<float>(ta, tb, n, n, k, (float)alpha, (const float *)dA, lda, (const float *)dA, lda,
                                 (float)beta, (float *)dC, ldc, up ? 2 : 1, s);
}

// Many right-hand sides (more than one sweep pass of 16): blocked over 64-row
// steps on the matrix cores -- x_i = Z_i r_i with Z_i the inverse of the
// diagonal block (forward; Z_i^T backward), then the rows still to be solved
// take the rank-64 update -L[>i, i] x_i (forward) or -L[i, <i]^T x_i
// (backward), each a GEMM.  R is rows x cols (ld ldr); Z holds nb 64 x 64
// inverses, Tmp 64 x cols.
template <typename T>
static int lx_b3_trsm_blocked(int rows, int cols, const T *Lt, int64_t na, T *R, int64_t ldr, int backward, T *Z,
                              T *Tmp, hipStream_t s)
{
    const int nb = (rows + 63) / 64, cus = lx_b3_cus();
    hipLaunchKernelGGL(lx_b3_inv_kernel<T>, dim3(nb), dim3(LX_WG_THREADS), 0, s, rows, Lt, na, Z);
    LX_HIP_TRY(hipGetLastError());
    for (int q = 0; q < nb; ++q) {
        const int i = backward ? nb - 1 - q : q, r0 = 64 * i, hb = std::min(64, rows - r0);
        T *Ri = R + r0;
        if (lx_gemm64<T>(backward, 0, hb, cols, hb, T(1), Z + (size_t)i * 4096, 64, Ri, ldr, T(0), Tmp, 64, 0,
                         (T *)nullptr, 0, cus, s))
            return LINOMPSS_ERR_HIP;
        LX_HIP_TRY(hipMemcpy2DAsync(Ri, sizeof(T) * ldr, Tmp, sizeof(T) * 64, sizeof(T) * hb, cols,
                                    hipMemcpyDeviceToDevice, s));
        if (!backward && r0 + hb < rows) {
            if (lx_gemm64<T>(0, 0, rows - r0 - hb, cols, hb, T(-1), Lt + (int64_t)r0 * na + r0 + hb, na, Ri, ldr,
                             T(1), Ri + hb, ldr, 0, (T *)nullptr, 0, cus, s))
                return LINOMPSS_ERR_HIP;
        } else if (backward && r0 > 0) {
            if (lx_gemm64<T>(1, 0, r0, cols, hb, T(-1), Lt + r0, na, Ri, ldr, T(1), R, ldr, 0, (T *)nullptr, 0, cus,
                             s))
                return LINOMPSS_ERR_HIP;
        }
    }
    return 0;
}

template <typename T>
static int lx_b3_trsm(int right, int upper, int trans, int unit, int m, int n, T alpha, const T *A, int64_t lda,
                      T *B, int64_t ldb, hipStream_t s)
{
    const int dtype = sizeof(T) == 8 ? 0 : 1;
    const int na = right ? n : m;  // order of A
    const int nrhs = right ? m : n;
    const bool blocked = nrhs > 16;
    const size_t lt_elems = (size_t)na * na, bt_elems = right ? (size_t)m * n : 0;
    const size_t z_elems = blocked ? (size_t)((na + 63) / 64) * 4096 : 0, tmp_elems = blocked ? (size_t)64 * nrhs : 0;
    T *Lt = (T *)lx_scratch(11, sizeof(T) * (lt_elems + bt_elems + z_elems + tmp_elems));
    if (!Lt) return LINOMPSS_ERR_NOMEM;
    hipLaunchKernelGGL(lx_b3_lower_kernel<T>, dim3(lx_b3_grid((int64_t)lt_elems)), dim3(256), 0, s, na, A, lda, upper,
                       unit, Lt);
    // left: op(A) X = alpha B.  op(A) = Lt (lower N, upper T) is a forward sweep,
    // op(A) = Lt^T (lower T, upper N) a backward one.  Right: X op(A) = alpha B
    // is op(A)^T X^T = alpha B^T, the left solve with the transpose flipped.
    const int t_left = right ? !trans : trans;
    const int which = (t_left != upper) ? 2 : 1;
    T *R = B;
    int64_t ldr = ldb;
    int rows = m, cols = n;
    if (right) {
        R = Lt + lt_elems;
        ldr = n;
        rows = n;
        cols = m;
        hipLaunchKernelGGL(lx_b3_transpose_kernel<T>, dim3((m + 31) / 32, (n + 31) / 32), dim3(256), 0, s, m, n, B, ldb,
                           R, ldr);
    }
    if (alpha != T(1))
        hipLaunchKernelGGL(lx_b3_scale_kernel<T>, dim3(lx_b3_grid((int64_t)rows * cols)), dim3(256), 0, s, rows, cols,
                           alpha, R, ldr);
    LX_HIP_TRY(hipGetLastError());
    if (alpha != T(0)) {
        const int rc = blocked ? lx_b3_trsm_blocked<T>(rows, cols, Lt, na, R, ldr, which == 2, Lt + lt_elems + bt_elems,
                                                       Lt + lt_elems + bt_elems + z_elems, s)
                               : lx_sweeps_dev(dtype, rows, cols, Lt, na, R, ldr, which, s);
        if (rc) return rc;
    }
    if (right) {
        hipLaunchKernelGGL(lx_b3_transpose_kernel<T>, dim3((n + 31) / 32, (m + 31) / 32), dim3(256), 0, s, n, m, R, ldr,
                           B, ldb);
        LX_HIP_TRY(hipGetLastError());
    }
    LX_HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

extern "C" int linompss_trsm_blas_dev(int dtype, char side, char uplo, char trans, char diag, int m, int n,
                                      double alpha, const void *dA, int64_t lda, void *dB, int64_t ldb, void *stream)
{
    const int sd = lx_b3_flag(side, "LR"), up = lx_b3_flag(uplo, "LU"), tr = lx_b3_flag(trans, "NT"),
              un = lx_b3_flag(diag, "NU");
    if (dtype != 0 && dtype != 1) return -1;
    if (sd < 0) return -2;
    if (up < 0) return -3;
    if (tr < 0) return -4;
    if (un < 0) return -5;
    if (m < 0) return -6;
    if (n < 0) return -7;
    if (lda < std::max(1, sd ? n : m)) return -10;
    if (ldb < std::max(1, m)) return -12;
    if (m == 0 || n == 0) return 0;
    std::lock_guard<std::mutex> g(lx_dev_mutex());
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
        return lx_b3_trsm<double>(sd, up, tr, un, m, n, alpha, (const double *)dA, lda, (double *)dB, ldb, s);
    return lx_b3_trsm<float>(sd, up, tr, un, m, n, (float)alpha, (const float *)dA, lda, (float *)dB, ldb, s);
}

// ---------------------------------------------------------------------------
// Host arrays, staged through device buffers (column-major, the caller's ld).

namespace {
template <typename T>
struct lx_b3_buf {
    T *d = nullptr;
    ~lx_b3_buf() { (void)hipFree(d); }
    bool alloc(size_t elems) { return hipMalloc(&d, sizeof(T) * std::max<size_t>(1, elems)) == hipSuccess; }
};
template <typename T>
int lx_b3_up(lx_b3_buf<T> &b, const T *h, int rows, int cols, int ld)
{
    if (!b.alloc((size_t)rows * cols)) return LINOMPSS_ERR_NOMEM;
    if (hipMemcpy2D(b.d, sizeof(T) * rows, h, sizeof(T) * (size_t)ld, sizeof(T) * rows, cols, hipMemcpyHostToDevice) !=
        hipSuccess)
        return LINOMPSS_ERR_HIP;
    return 0;
}
template <typename T>
int lx_b3_down(T *h, int ld, const lx_b3_buf<T> &b, int rows, int cols)
{
    if (hipMemcpy2D(h, sizeof(T) * (size_t)ld, b.d, sizeof(T) * rows, sizeof(T) * rows, cols, hipMemcpyDeviceToHost) !=
        hipSuccess)
        return LINOMPSS_ERR_HIP;
    return 0;
}

template <typename T>
int host_gemm(char transa, char transb, int b, int d, int c, int m, int n, int k, T alpha, const T *A, int lda,
              const T *B, int ldb, T beta, T *C, int ldc)
{
    const int ta = lx_b3_flag(transa, "NT"), tb = lx_b3_flag(transb, "NT");
    if (ta < 0) return -1;
    if (tb < 0) return -2;
    if (b <= 0) return -3;
    if (d <= 0) return -4;
    if (c <= 0) return -5;
    if (m < 0) return -6;
    if (n < 0) return -7;
    if (k < 0) return -8;
    const int ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
    if (lda < std::max(1, ar)) return -11;
    if (ldb < std::max(1, br)) return -13;
    if (ldc < std::max(1, m)) return -16;
    if (m == 0 || n == 0) return 0;
    lx_b3_buf<T> dA, dB, dC;
    int rc = 0;
    if (k > 0 && ((rc = lx_b3_up(dA, A, ar, ac, lda)) || (rc = lx_b3_up(dB, B, br, bc, ldb)))) return rc;
    if ((rc = lx_b3_up(dC, C, m, n, ldc))) return rc;
    rc = linompss_gemm_dev(sizeof(T) == 8 ? 0 : 1, transa, transb, m, n, k, alpha, dA.d, std::max(1, ar), dB.d,
                           std::max(1, br), beta, dC.d, m, nullptr);
    return rc ? rc : lx_b3_down(C, ldc, dC, m, n);
}

template <typename T>
int host_syrk(char uplo, char trans, int b, int n, int k, T alpha, const T *A, int lda, T beta, T *C, int ldc)
{
    const int up = lx_b3_flag(uplo, "LU"), tr = lx_b3_flag(trans, "NT");
    if (up < 0) return -1;
    if (tr < 0) return -2;
    if (b <= 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    const int ar = tr ? k : n, ac = tr ? n : k;
    if (lda < std::max(1, ar)) return -8;
    if (ldc < std::max(1, n)) return -11;
    if (n == 0) return 0;
    lx_b3_buf<T> dA, dC;
    int rc = 0;
    if (k > 0 && (rc = lx_b3_up(dA, A, ar, ac, lda))) return rc;
    if ((rc = lx_b3_up(dC, C, n, n, ldc))) return rc;
    rc = linompss_syrk_dev(sizeof(T) == 8 ? 0 : 1, uplo, trans, n, k, alpha, dA.d, std::max(1, ar), beta, dC.d, n,
                           nullptr);
    return rc ? rc : lx_b3_down(C, ldc, dC, n, n);
}

template <typename T>
int host_trsm(char side, char uplo, char trans, char diag, int m, int n, int b, T alpha, const T *A, int lda, T *B,
              int ldb)
{
    const int sd = lx_b3_flag(side, "LR");
    if (sd < 0) return -1;
    if (lx_b3_flag(uplo, "LU") < 0) return -2;
    if (lx_b3_flag(trans, "NT") < 0) return -3;
    if (lx_b3_flag(diag, "NU") < 0) return -4;
    if (m < 0) return -5;
    if (n < 0) return -6;
    if (b <= 0) return -7;
    const int na = sd ? n : m;
    if (lda < std::max(1, na)) return -10;
    if (ldb < std::max(1, m)) return -12;
    if (m == 0 || n == 0) return 0;
    lx_b3_buf<T> dA, dB;
    int rc = 0;
    if ((rc = lx_b3_up(dA, A, na, na, lda)) || (rc = lx_b3_up(dB, B, m, n, ldb))) return rc;
    rc = linompss_trsm_blas_dev(sizeof(T) == 8 ? 0 : 1, side, uplo, trans, diag, m, n, alpha, dA.d, na, dB.d, m,
                                nullptr);
    return rc ? rc : lx_b3_down(B, ldb, dB, m, n);
}
}  // namespace

extern "C" int linompss_dgemm(char transa, char transb, int b, int d, int c, int m, int n, int k, double alpha,
                              const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
    return host_gemm<double>(transa, transb, b, d, c, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
extern "C" int linompss_sgemm(char transa, char transb, int b, int d, int c, int m, int n, int k, float alpha,
                              const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc)
{
    return host_gemm<float>(transa, transb, b, d, c, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
extern "C" int linompss_dsyrk(char uplo, char trans, int b, int n, int k, double alpha, const double *A, int lda,
                              double beta, double *C, int ldc)
{
    return host_syrk<double>(uplo, trans, b, n, k, alpha, A, lda, beta, C, ldc);
}
extern "C" int linompss_ssyrk(char uplo, char trans, int b, int n, int k, float alpha, const float *A, int lda,
                              float beta, float *C, int ldc)
{
    return host_syrk<float>(uplo, trans, b, n, k, alpha, A, lda, beta, C, ldc);
}
extern "C" int linompss_dtrsm(char side, char uplo, char trans, char diag, int m, int n, int b, double alpha,
                              const double *A, int lda, double *B, int ldb)
{
    return host_trsm<double>(side, uplo, trans, diag, m, n, b, alpha, A, lda, B, ldb);
}
extern "C" int linompss_strsm(char side, char uplo, char trans, char diag, int m, int n, int b, float alpha,
                              const float *A, int lda, float *B, int ldb)
{
    return host_trsm<float>(side, uplo, trans, diag, m, n, b, alpha, A, lda, B, ldb);
}
