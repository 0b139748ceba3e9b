// lx_adjoint.hip -- reverse-mode (adjoint) kernels of the FP64 Cholesky
// factor and of the triangular solves.
//
// Cholesky adjoint (blocked reverse mode of the left-looking block Cholesky,
// I. Murray, "Differentiation of the Cholesky decomposition", 2016): with
// A = L L^T read from its lower triangle and Lbar the adjoint of L (lower),
// Abar (lower representation: dLoss = sum_{i>=j} Abar_ij dA_ij) is computed in
// place of Lbar.  Forward block step for block column J (rows j0..j0+jb):
//     D = chol(A_D - R R^T),   C = (A_C - B R^T) D^-T
// with R = L[J, <J], B = L[>J, <J], C = L[>J, J].  Reverse, J from last to first:
//     Qbar = Cbar D^-1              (= Abar_C)
//     Dbar -= tril(Qbar^T C)
//     Pbar  = level-2 adjoint of D = chol(P) on the diagonal block (= Abar_D)
//     Bbar -= Qbar R
//     Rbar -= Qbar^T B + (Pbar + Pbar^T) R
// The forward factor has no counterpart in the reference beyond task_dpotrf
// (tasks_potrf.c:20-24); the adjoint is new.  Work ~ 2 n^3 / 3 flops on the
// FP64 matrix cores through one tiled GEMM kernel (op(A), op(B) staged through
// LDS, optional deterministic split-K), plus one workgroup per block for the
// diagonal steps.
//
// Triangular-solve adjoint: x = op(L)^-1 b with op = N or T; given xbar,
//     bbar = op(L)^-T xbar,  Lbar -= tril(bbar x^T)  (op = N)
//                            Lbar -= tril(x bbar^T)  (op = T).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lx_device.hpp"
#ifdef LX_ADJOINT_KERNELS_ONLY
#define LX_GEMM64_KERNELS_ONLY
#endif
#include "lx_gemm64.hpp"

// Diagonal block step of the reverse sweep:
//   mode 0: the inverses of all diagonal blocks of an n-row lower triangle
//           (n passed as jb), one workgroup per 64-column block J:
//           Z_J = L[J, J]^-1 into Zb + 4096 J (64 x 64, ld 64; identity past
//           the block's order);
//   mode 1: Pbar = the adjoint of D = chol(P) applied to Dbar (Abar[J, J],
//           lower), written back in place (lower representation); the
//           leading jb x jb part of Sb (ld lds, which may be jb) =
//           sscale (Pbar + Pbar^T), nothing past it.  Zb holds Z = D^-1 from
//           mode 0.
// Mode 1 is Murray's closed form for one block (eq. 10 of the paper in the
// file header), three 64 x 64 x 64 products on the matrix cores:
//     M = D^T Dbar,  Phi = tril(M) with its diagonal halved,
//     G = Z^T Phi Z,  Pbar_ij = G_ij + G_ji (i > j),  Pbar_ii = G_ii,
// the same result as the column-by-column (level-2) recurrences
//     cbar = Dbar[>j, j] / d_j; Dbar[j, j] -= cbar . c; pbar = Dbar[j, j] / (2 d_j);
//     Dbar[>j, <j] -= cbar L[j, <j]; Dbar[j, <j] -= cbar^T L[>j, <j] + 2 pbar L[j, <j]
// (which took 64 dependent steps of three barriers each, about 120 us per block).
// A block of order jb < 64 is padded: D and Z with the identity, Dbar with
// zeros, so every product keeps the padding out of the leading jb x jb part.
__device__ __forceinline__ void lx_adj_mm64(lx_d4 (&acc)[4], const double (*A)[65], bool at, const double (*B)[65])
{
    // the wave's 16-row strip of op(A) B (op(A) = A^T when at), 64 x 64, from LDS
    const int lane = lx_lane(), lr = lane & 15, lk = lane >> 4, r0 = 16 * lx_wave();
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = lx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < 64; k0 += 4) {
        const double a = at ? A[k0 + lk][r0 + lr] : A[r0 + lr][k0 + lk];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = lx_mfma<double>::mac(a, B[k0 + lk][16 * j + lr], acc[j]);
    }
}

__global__ void __launch_bounds__(LX_WG_THREADS)
lx_adj_diag_kernel(int mode, int jb, const double *__restrict__ Ld, int64_t ldl, double *__restrict__ Ab,
                   int64_t lda, double *__restrict__ Zb, double *__restrict__ Sb, int64_t lds, double sscale)
{
    __shared__ double sL[64][65], sX[64][65];
    if (mode == 0) {
        const int j0 = 64 * (int)blockIdx.x;
        jb = min(64, jb - j0);
        Ld += (int64_t)j0 * ldl + j0;
        Zb += (int64_t)blockIdx.x * 4096;
    }
    const int tid = lx_tid(), lane = tid & 63, lr = lane & 15, lk = lane >> 4, r0 = 16 * lx_wave();
    // the blocks' loads (and, in mode 1, Z's) all issued before the LDS stores
    double vl[16], vx[16], vz[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = tid + i * LX_WG_THREADS, r = e & 63, c = e >> 6;
        const bool in = r < jb && c < jb && r >= c;
        vl[i] = in ? Ld[(int64_t)c * ldl + r] : (r == c ? 1.0 : 0.0);
        vx[i] = (mode == 1 && in) ? Ab[(int64_t)c * lda + r] : 0.0;
        vz[i] = mode == 1 ? Zb[e] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = tid + i * LX_WG_THREADS, r = e & 63, c = e >> 6;
        sL[r][c] = vl[i];
        sX[r][c] = vx[i];
    }
    __syncthreads();
    if (mode == 0) {
        lx_tri_inv64<double>(sL, sX);
        for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
            const int r = e & 63, c = e >> 6;
            Zb[c * 64 + r] = sX[r][c];
        }
        return;
    }
    lx_d4 acc[4];
    // M = D^T Dbar, then Phi into sX and Z into sL
    lx_adj_mm64(acc, sL, true, sX);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + lx_mfma<double>::row(lk, e), c = 16 * j + lr;
            sX[r][c] = r > c ? acc[j][e] : (r == c ? 0.5 * acc[j][e] : 0.0);
        }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = tid + i * LX_WG_THREADS;
        sL[e & 63][e >> 6] = vz[i];
    }
    __syncthreads();
    // T = Z^T Phi into sX
    lx_adj_mm64(acc, sL, true, sX);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) sX[r0 + lx_mfma<double>::row(lk, e)][16 * j + lr] = acc[j][e];
    __syncthreads();
    // G = T Z into sX
    lx_adj_mm64(acc, sX, false, sL);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) sX[r0 + lx_mfma<double>::row(lk, e)][16 * j + lr] = acc[j][e];
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
        const int r = e & 63, c = e >> 6;
        const double g = sX[r][c] + sX[c][r];
        if (r < jb && c < jb && r >= c) Ab[(int64_t)c * lda + r] = r > c ? g : sX[r][r];
        if (r < jb && c < jb) Sb[c * lds + r] = sscale * g;
    }
}

// Lbar_ii += 2 w / L_ii (the adjoint of w log det A = 2 w sum_i log L_ii).
__global__ void lx_adj_logdet_kernel(int n, const double *__restrict__ L, int64_t ldl, double *__restrict__ Lb,
                                     int64_t ldlb, double w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Lb[(int64_t)i * ldlb + i] += 2.0 * w / L[(int64_t)i * ldl + i];
}

// out[0] = 2 sum_i log L_ii (log det of A = L L^T), one workgroup, fixed order.
__global__ void __launch_bounds__(LX_WG_THREADS) lx_adj_logdet_sum_kernel(int n, const double *__restrict__ L,
                                                                          int64_t ldl, double *__restrict__ out)
{
    __shared__ double red[LX_WG_THREADS];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += LX_WG_THREADS) s += log(L[(int64_t)i * ldl + i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = LX_WG_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = 2.0 * red[0];
}

// The group's Qbar blocks into Abar: column c of Q (rows x cols, ld ldq; the
// group's first row and column at Q[0]) holds Qbar of block c / 64 in the
// rows below that block, which are Abar's final values there.
__global__ void lx_adj_qcopy_kernel(int rows, int cols, const double *__restrict__ Q, int64_t ldq,
                                    double *__restrict__ Ab, int64_t lda)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int r = (int)(e % rows), c = (int)(e / rows);
    if (r >= (c / 64) * 64 + 64) Ab[(int64_t)c * lda + r] = Q[(int64_t)c * ldq + r];
}

// Y (ld) := X, a rows x cols block copy.
__global__ void lx_adj_copy_kernel(int rows, int cols, const double *__restrict__ X, int64_t ldx,
                                   double *__restrict__ Y, int64_t ldy)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int r = (int)(e % rows), c = (int)(e / rows);
    Y[(int64_t)c * ldy + r] = X[(int64_t)c * ldx + r];
}

#ifndef LX_ADJOINT_KERNELS_ONLY
#include <algorithm>
#include <mutex>

#include "lx_internal.hpp"

namespace {

std::mutex lx_adj_mu;  // the adjoint's scratch slot

int lx_adj_cus()
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

// C = alpha op(A) op(B) + beta C (lx_gemm64, FP64), split-K through part.
// The sweep's skinny products (64 x 64 or 64 x 192 outputs over thousands of
// rows) split K into chunks of 64: at 256 a chunk's workgroup walked its 16
// slabs one load round trip at a time (about 20 us).
int lx_adj_gemm(int ta, int tb, int M, int N, int K, double alpha, const double *A, int64_t lda, const double *B,
                int64_t ldb, double beta, double *C, int64_t ldc, int lower, double *part, size_t part_elems,
                hipStream_t s)
{
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    return lx_gemm64<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower ? 1 : 0, part, part_elems,
                             lx_adj_cus(), s, 64);
}

__global__ void lx_adj_zero_kernel(double *__restrict__ X, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) X[e] = 0.0;
}

// The reverse sweep over 64-column blocks (see the header comment), in groups
// of four blocks (256 columns, from the last group to the first).  Inside a
// group each block runs the recurrences with its updates restricted to the
// group's columns left of it; the updates of the columns left of the group
// are gathered and applied once per group as two rank-256 products.  With
// Q_G the group's blocks side by side (block J's column block holds S_J / 2
// in block J's own rows, S_J = Pbar_J + Pbar_J^T, and Qbar_J in the rows
// below; zeros above), rows from the group's first row g0 to n, and
// L_G = L[group rows, 0:g0]:
//     Abar[g0:n, 0:g0]     -= Q_G L_G                (the Bbar terms, and S R / 2)
//     Abar[group, 0:g0]    -= Q_G^T L[g0:n, 0:g0]    (the Qbar^T B terms, and S^T R / 2)
// S is symmetric, so the two halves add up to the S R term of the block's
// left columns; inside the group the same column block of Q_G, from block
// J's first row, gives the in-group terms of both kinds in two products.  The
// same sums as the block-by-block sweep in another order: the rank-64
// updates of the whole left part, each reading and writing it, become two
// rank-256 ones, and a block's chain of launches carries no S R product.
// Abar_C = Qbar_J is final once formed, so the group's blocks are copied into
// Abar once, at the group's end.
int lx_chol_adjoint(int n, const double *L, int64_t ldl, double *Ab, int64_t lda, hipStream_t s)
{
    const int nb = (n + 63) / 64;
    constexpr int GB = 4;  // blocks per group
    const size_t part_elems = (size_t)16 * 64 * std::max(n, 64);
    const size_t qg_elems = (size_t)std::max(n, 1) * 64 * GB;
    double *ws = (double *)lx_scratch(
        9, sizeof(double) * ((size_t)nb * 4096 + 64 * 64 + qg_elems + part_elems));
    if (!ws) return LINOMPSS_ERR_NOMEM;
    double *Zall = ws, *QG = Zall + (size_t)nb * 4096 + 64 * 64, *part = QG + qg_elems;
    // the inverses of the diagonal blocks do not depend on the sweep: all at once
    hipLaunchKernelGGL(lx_adj_diag_kernel, dim3(nb), dim3(LX_WG_THREADS), 0, s, 0, n, L, ldl, Ab, lda, Zall,
                       nullptr, (int64_t)64, 0.0);
    int rc = 0;
    for (int G = (nb - 1) / GB; G >= 0 && rc == 0; --G) {
        const int gJ0 = G * GB, gJ1 = std::min(nb, gJ0 + GB) - 1;  // blocks of the group
        const int g0 = 64 * gJ0, gw = std::min(n, 64 * (gJ1 + 1)) - g0;
        const int64_t ldqg = std::max(n - g0, 1);
        const int64_t qtot = ldqg * gw;
        hipLaunchKernelGGL(lx_adj_zero_kernel, dim3((unsigned)((qtot + 255) / 256)), dim3(256), 0, s, QG, qtot);
        for (int J = gJ1; J >= gJ0 && rc == 0; --J) {
            const int j0 = 64 * J, jb = std::min(64, n - j0), m = n - j0 - jb;
            const int pin = j0 - g0;  // the group's columns left of block J
            const double *D = L + (int64_t)j0 * ldl + j0;
            double *Dbar = Ab + (int64_t)j0 * lda + j0;
            // block J's column block of Q_G from its first row: S_J / 2, then Qbar_J
            double *Qs = QG + (int64_t)(j0 - g0) * ldqg + (j0 - g0);
            double *Qb = Qs + jb;
            const double *Zb = Zall + (size_t)J * 4096;
            if (m > 0) {
                const double *Cbar = Ab + (int64_t)j0 * lda + j0 + jb;
                const double *C = L + (int64_t)j0 * ldl + j0 + jb;
                // Qbar = Cbar D^-1 (Abar_C = Qbar, copied at the group's end)
                if ((rc = lx_adj_gemm(0, 0, m, jb, jb, 1.0, Cbar, lda, Zb, 64, 0.0, Qb, ldqg, 0, part, part_elems, s)))
                    break;
                // Dbar -= tril(Qbar^T C)
                if ((rc = lx_adj_gemm(1, 0, jb, jb, m, -1.0, Qb, ldqg, C, ldl, 1.0, Dbar, lda, 1, part, part_elems,
                                      s)))
                    break;
            }
            // Pbar (in place of Dbar) and S_J / 2 into Q_G
            hipLaunchKernelGGL(lx_adj_diag_kernel, dim3(1), dim3(LX_WG_THREADS), 0, s, 1, jb, D, ldl, Dbar, lda,
                               Zall + (size_t)J * 4096, Qs, ldqg, 0.5);
            if (pin > 0) {
                // inside the group (columns g0 .. j0), rows from block J on:
                //   Abar[j0:n, .]    -= [S/2; Qbar] R         (Bbar and S R / 2)
                //   Abar[J, .]       -= [S/2; Qbar]^T L[j0:n] (Qbar^T B and S^T R / 2)
                const double *Rin = L + (int64_t)g0 * ldl + j0;
                if ((rc = lx_adj_gemm(0, 0, m + jb, pin, jb, -1.0, Qs, ldqg, Rin, ldl, 1.0, Ab + (int64_t)g0 * lda + j0,
                                      lda, 0, part, part_elems, s)))
                    break;
                if ((rc = lx_adj_gemm(1, 0, jb, pin, m + jb, -1.0, Qs, ldqg, Rin, ldl, 1.0, Ab + (int64_t)g0 * lda + j0,
                                      lda, 0, part, part_elems, s)))
                    break;
            }
        }
        if (rc) break;
        // the group's Qbar blocks into Abar (rows below each block)
        hipLaunchKernelGGL(lx_adj_qcopy_kernel, dim3((unsigned)((qtot + 255) / 256)), dim3(256), 0, s, (int)ldqg, gw,
                           QG, ldqg, Ab + (int64_t)g0 * lda + g0, lda);
        if (g0 > 0) {
            // the columns left of the group: the rank-gw products with Q_G
            if ((rc = lx_adj_gemm(0, 0, n - g0, g0, gw, -1.0, QG, ldqg, L + g0, ldl, 1.0, Ab + g0, lda, 0, part,
                                  part_elems, s)))
                break;
            if ((rc = lx_adj_gemm(1, 0, gw, g0, n - g0, -1.0, QG, ldqg, L + g0, ldl, 1.0, Ab + g0, lda, 0, part,
                                  part_elems, s)))
                break;
        }
    }
    if (rc == 0) LX_HIP_TRY(hipGetLastError());
    return rc;
}

// Lbar = beta Lbar - tril(U V^T) for n x nrhs U, V (the rank-nrhs term of
// the solve adjoint); beta 0 writes the lower triangle without reading it.
int lx_trsm_adjoint_update(int n, int nrhs, const double *U, int64_t ldu, const double *V, int64_t ldv,
                           double *Lb, int64_t ldlb, double beta, hipStream_t s)
{
    return lx_adj_gemm(0, 1, n, n, nrhs, -1.0, U, ldu, V, ldv, beta, Lb, ldlb, 1, nullptr, 0, s);
}

}  // namespace

extern "C" int linompss_dpotrf_adjoint_dev(int n, const double *dL, int64_t ldl, double *dLbar, int64_t ldlbar,
                                           double logdet_weight, void *stream)
{
    if (n < 0) return -1;
    if (ldl < std::max(1, n)) return -3;
    if (ldlbar < std::max(1, n)) return -5;
    if (n == 0) return 0;
    std::lock_guard<std::mutex> g(lx_adj_mu);
    hipStream_t s = (hipStream_t)stream;
    if (logdet_weight != 0.0)
        hipLaunchKernelGGL(lx_adj_logdet_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, dL, ldl, dLbar, ldlbar,
                           logdet_weight);
    const int rc = lx_chol_adjoint(n, dL, ldl, dLbar, ldlbar, s);
    if (rc) return rc;
    LX_HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

extern "C" int linompss_dtrsm_adjoint_dev(int trans, int n, int nrhs, const double *dL, int64_t ldl,
                                          const double *dX, int64_t ldx, double *dXbar, int64_t ldxbar,
                                          double *dLbar, int64_t ldlbar, int accumulate, void *stream)
{
    if (trans != 0 && trans != 1) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (ldl < std::max(1, n)) return -5;
    if (ldx < std::max(1, n)) return -7;
    if (ldxbar < std::max(1, n)) return -9;
    if (dLbar && ldlbar < std::max(1, n)) return -11;
    if (n == 0 || nrhs == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    {
        // bbar = op(L)^-T xbar: the backward sweep for op = N, the forward one for op = T
        std::lock_guard<std::mutex> g(lx_dev_mutex());
        rc = lx_sweeps_dev(0, n, nrhs, dL, ldl, dXbar, ldxbar, trans ? 1 : 2, s);
    }
    if (rc) return rc;
    if (dLbar) {
        std::lock_guard<std::mutex> g(lx_adj_mu);
        const double beta = accumulate ? 1.0 : 0.0;
        rc = trans ? lx_trsm_adjoint_update(n, nrhs, dX, ldx, dXbar, ldxbar, dLbar, ldlbar, beta, s)
                   : lx_trsm_adjoint_update(n, nrhs, dXbar, ldxbar, dX, ldx, dLbar, ldlbar, beta, s);
        if (rc) return rc;
    }
    LX_HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

extern "C" int linompss_dlogdet_dev(int n, const double *dL, int64_t ldl, double *logdet, void *stream)
{
    if (n < 0) return -1;
    if (ldl < std::max(1, n)) return -3;
    if (!logdet) return -4;
    *logdet = 0.0;
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(lx_adj_mu);
    double *ws = (double *)lx_scratch(10, 256);
    if (!ws) return LINOMPSS_ERR_NOMEM;
    hipLaunchKernelGGL(lx_adj_logdet_sum_kernel, dim3(1), dim3(LX_WG_THREADS), 0, s, n, dL, ldl, ws);
    LX_HIP_TRY(hipGetLastError());
    LX_HIP_TRY(hipMemcpyAsync(logdet, ws, sizeof(double), hipMemcpyDeviceToHost, s));
    LX_HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// Host arrays: Abar from A itself -- the factor is recomputed on the device
// (tiled FP64 Cholesky, tile b) rather than kept from the forward pass.
extern "C" int linompss_dpotrf_adjoint(int n, int b, const double *A, int lda, double *Lbar, int ldlbar,
                                       double logdet_weight)
{
    if (n < 0) return -1;
    if (b <= 0) return -2;
    if (lda < std::max(1, n)) return -4;
    if (ldlbar < std::max(1, n)) return -6;
    if (n == 0) return 0;
    double *dA = nullptr, *dB = nullptr;
    const size_t nn = (size_t)n * n;
    if (hipMalloc(&dA, sizeof(double) * nn) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    if (hipMalloc(&dB, sizeof(double) * nn) != hipSuccess) {
        (void)hipFree(dA);
        return LINOMPSS_ERR_NOMEM;
    }
    int rc = 0;
    if (hipMemcpy2D(dA, sizeof(double) * n, A, sizeof(double) * (size_t)lda, sizeof(double) * n, n,
                    hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy2D(dB, sizeof(double) * n, Lbar, sizeof(double) * (size_t)ldlbar, sizeof(double) * n, n,
                    hipMemcpyHostToDevice) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc) rc = linompss_chol_dev(0, 0, n, b, dA, n, nullptr, nullptr);
    if (!rc) rc = linompss_dpotrf_adjoint_dev(n, dA, n, dB, n, logdet_weight, nullptr);
    if (!rc && hipMemcpy2D(Lbar, sizeof(double) * (size_t)ldlbar, dB, sizeof(double) * n, sizeof(double) * n, n,
                           hipMemcpyDeviceToHost) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    (void)hipFree(dA);
    (void)hipFree(dB);
    return rc;
}

// Host arrays: x = op(L)^-1 b was computed with L (lower, ldl); xbar (in) ->
// bbar (out), Lbar -= the rank-nrhs term (Lbar may be NULL).
extern "C" int linompss_dtrsm_adjoint(char trans, int n, int nrhs, const double *L, int ldl, const double *X,
                                      int ldx, double *Xbar, int ldxbar, double *Lbar, int ldlbar)
{
    const int tr = (trans == 'T' || trans == 't') ? 1 : ((trans == 'N' || trans == 'n') ? 0 : -1);
    if (tr < 0) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (ldl < std::max(1, n)) return -5;
    if (ldx < std::max(1, n)) return -7;
    if (ldxbar < std::max(1, n)) return -9;
    if (Lbar && ldlbar < std::max(1, n)) return -11;
    if (n == 0 || nrhs == 0) return 0;
    const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * (nn * (Lbar ? 2 : 1) + 2 * nr)) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    double *dL = d, *dX = d + nn, *dXb = dX + nr, *dLb = Lbar ? dXb + nr : nullptr;
    auto up = [&](double *dst, const double *src, int ld, int cols) {
        return hipMemcpy2D(dst, sizeof(double) * n, src, sizeof(double) * (size_t)ld, sizeof(double) * n, cols,
                           hipMemcpyHostToDevice) == hipSuccess;
    };
    int rc = 0;
    if (!up(dL, L, ldl, n) || !up(dX, X, ldx, nrhs) || !up(dXb, Xbar, ldxbar, nrhs) ||
        (Lbar && !up(dLb, Lbar, ldlbar, n)))
        rc = LINOMPSS_ERR_HIP;
    if (!rc) rc = linompss_dtrsm_adjoint_dev(tr, n, nrhs, dL, n, dX, n, dXb, n, dLb, n, 1, nullptr);
    if (!rc && hipMemcpy2D(Xbar, sizeof(double) * (size_t)ldxbar, dXb, sizeof(double) * n, sizeof(double) * n, nrhs,
                           hipMemcpyDeviceToHost) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc && Lbar &&
        hipMemcpy2D(Lbar, sizeof(double) * (size_t)ldlbar, dLb, sizeof(double) * n, sizeof(double) * n, n,
                    hipMemcpyDeviceToHost) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    (void)hipFree(d);
    return rc;
}
#endif  // LX_ADJOINT_KERNELS_ONLY
