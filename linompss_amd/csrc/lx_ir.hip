// lx_ir.hip -- mixed-precision iterative refinement on the device.
//
// Loop contract of the reference's itref (src/itref/itref.c:999-1101):
//   x0 = solve(b); r = b - A x (FP64, itref.c:1010-1020)
//   for i < refit: test ||r||; d = solve(r) (itref.c:1062); x += d
//                  (bblas_sdcpaxpy, itref.c:1084); r = b - A x (itref.c:1094)
// Here solve() is the FP32 tiled Cholesky factor + sweeps of this library,
// and the FP64 pieces (residual, update, backward-error test) are the kernels
// below.  A is symmetric and only its lower triangle is read: the residual is
// r = b - (tril(A) x + stril(A)^T x), one pass over the lower triangle that
// also produces ||A||_inf on the first refinement step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <mutex>
#include <vector>

#include "linompss.h"
#include "lx_device.hpp"
#include "lx_internal.hpp"


// Residual pass over the lower triangle, read once.  The triangle is cut into
// 256 x 256 super-blocks (4 x 4 blocks of 64); the workgroup of super-block
// (SI, SJ), SJ <= SI, forms for each of its blocks A_IJ both
//   the row part    A_IJ x_J      (contributes to y_I)
//   the column part A_IJ^T x_I    (contributes to y_J; on a diagonal block the
//                                  strictly lower part only)
// and, when ABS, the matching sums of |a_ij| (the rows of |A| for ||A||_inf).
// Its sums over the super-block go to P[SI][SJ] (row parts) and P[SJ][SI]
// (column parts; both parts of a diagonal super-block go to P[SI][SI]), and
// y_S = sum over SJ of P[S][SJ] in a fixed order (lx_symv_reduce_kernel):
// deterministic, no atomics.  Each 64 x 64 block is staged in LDS (coalesced
// column reads); thread (l, q) owns row l / column l of the block and the
// quarter q of the other index.  P holds ns*ns slots of 256 x NC values
// (NC = NR, plus 1 for the |A| sums).
#define LX_SY_BS 64
#define LX_SY_SB 4  // blocks per super-block side
template <int NR, bool ABS>
__global__ void __launch_bounds__(256) lx_symv_blk_kernel(const double *__restrict__ A, int64_t lda, int n,
                                                          const double *__restrict__ X, int64_t ldx, int nrhs,
                                                          double *__restrict__ P)
{
    constexpr int NC = NR + (ABS ? 1 : 0), SB = LX_SY_SB, SN = SB * LX_SY_BS;
    __shared__ double sA[64][65];           // sA[c][r] = A[I*64 + r, J*64 + c]
    __shared__ double sxI[SB][NR][64];      // x of the super-block's row blocks
    __shared__ double sxJ[SB][NR][64];      // x of its column blocks
    __shared__ double sred[4][NC][65];
    const int nb = (n + 63) / 64, ns = (nb + SB - 1) / SB;
    const int64_t nsb = (int64_t)ns * (ns + 1) / 2;
    const int tid = threadIdx.x, l = tid & 63, q = tid >> 6;
    for (int64_t sbk = blockIdx.x; sbk < nsb; sbk += gridDim.x) {
        int SI = (int)((sqrt(8.0 * (double)sbk + 1.0) - 1.0) * 0.5);
        while ((int64_t)SI * (SI + 1) / 2 > sbk) --SI;
        while ((int64_t)(SI + 1) * (SI + 2) / 2 <= sbk) ++SI;
        const int SJ = (int)(sbk - (int64_t)SI * (SI + 1) / 2);
        __syncthreads();  // the previous super-block is done with the LDS
        for (int e = tid; e < SB * NR * 64; e += 256) {
            const int k = e % 64, c = (e / 64) % NR, b = e / (64 * NR);
            const int ri = SI * SN + b * 64 + k, rj = SJ * SN + b * 64 + k;
            sxI[b][c][k] = (c < nrhs && ri < n) ? X[(int64_t)c * ldx + ri] : 0.0;
            sxJ[b][c][k] = (c < nrhs && rj < n) ? X[(int64_t)c * ldx + rj] : 0.0;
        }
        double accR[SB][NC], accC[SB][NC];
#pragma unroll
        for (int b = 0; b < SB; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) accR[b][c] = accC[b][c] = 0.0;
        // blocks t = bi * SB + bj of the super-block in order; the 16 values a
        // thread stages of the next one are loaded while the current one is
        // multiplied
        auto blk_ok = [&](int t) {
            const int I = SI * SB + t / SB, J = SJ * SB + t % SB;
            return J <= I && I < nb;
        };
        double nxt[16];
        auto gload = [&](int t) {
            const int I = SI * SB + t / SB, J = SJ * SB + t % SB;
            const int r0 = I * 64, c0 = J * 64, hr = min(64, n - r0), hc = min(64, n - c0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int e = tid + 256 * i, r = e & 63, c = e >> 6;
                nxt[i] = (r < hr && c < hc && (I != J || r >= c)) ? A[(int64_t)(c0 + c) * lda + r0 + r] : 0.0;
            }
        };
        {
            int t = 0;
            while (t < SB * SB && !blk_ok(t)) ++t;  // uniform
            if (t < SB * SB) gload(t);
        }
#pragma unroll
        for (int bi = 0; bi < SB; ++bi) {
#pragma unroll
            for (int bj = 0; bj < SB; ++bj) {
                const int I = SI * SB + bi, J = SJ * SB + bj;
                if (J > I || I >= nb) continue;  // uniform
                __syncthreads();  // sA is free; also orders the sx writes above
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int e = tid + 256 * i;
                    sA[e >> 6][e & 63] = nxt[i];
                }
                __syncthreads();
                {
                    int t = bi * SB + bj + 1;
                    while (t < SB * SB && !blk_ok(t)) ++t;  // uniform
                    if (t < SB * SB) gload(t);
                }
                // row part: row l, columns 16q .. 16q+15
#pragma unroll 4
                for (int k = 16 * q; k < 16 * q + 16; ++k) {
                    const double a = sA[k][l];
#pragma unroll
                    for (int c = 0; c < NR; ++c) accR[bi][c] += a * sxJ[bj][c][k];
                    if (ABS) accR[bi][NR] += fabs(a);
                }
                // column part: column l, rows 16q .. 16q+15 (diagonal block: rows > l only)
#pragma unroll 4
                for (int r = 16 * q; r < 16 * q + 16; ++r) {
                    const double a = (I != J || r > l) ? sA[l][r] : 0.0;
#pragma unroll
                    for (int c = 0; c < NR; ++c) accC[bj][c] += a * sxI[bi][c][r];
                    if (ABS) accC[bj][NR] += fabs(a);
                }
            }
        }
        // quarter sums in a fixed order, then out to P
        double *Prow = P + ((size_t)SI * ns + SJ) * SN * NC;
        double *Pcol = P + ((size_t)SJ * ns + SI) * SN * NC;
        for (int b = 0; b < SB; ++b) {
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NC; ++c) sred[q][c][l] = accR[b][c];
            __syncthreads();
            if (SI == SJ) {  // uniform: both parts land in P[SI][SI]
#pragma unroll
                for (int c = 0; c < NC; ++c) sred[q][c][l] += accC[b][c];
                __syncthreads();
            }
            for (int e = tid; e < NC * 64; e += 256) {
                const int r = e % 64, c = e / 64;
                Prow[(size_t)c * SN + b * 64 + r] = (sred[0][c][r] + sred[1][c][r]) + (sred[2][c][r] + sred[3][c][r]);
            }
            if (SI != SJ) {
                __syncthreads();
#pragma unroll
                for (int c = 0; c < NC; ++c) sred[q][c][l] = accC[b][c];
                __syncthreads();
                for (int e = tid; e < NC * 64; e += 256) {
                    const int r = e % 64, c = e / 64;
                    Pcol[(size_t)c * SN + b * 64 + r] =
                        (sred[0][c][r] + sred[1][c][r]) + (sred[2][c][r] + sred[3][c][r]);
                }
            }
        }
    }
}

// The same residual pass for up to 16 right-hand sides at once, the block
// products on the matrix cores (v_mfma_f64_16x16x4f64; lane l supplies
// A[l%16][l/16] and B[l/16][l%16], value i of the result is row l/16 + 4i of
// column l%16).  Per 64 x 64 block, wave q forms the 16 x 16 tiles
//   rows 16q.. of A_IJ x_J       (row part; K = the block's 64 columns)
//   rows 16q.. of A_IJ^T x_I     (column part; strictly lower on a diagonal block)
// with A from the block staged in LDS and x from LDS ([b][column][row], row
// stride 66); the |A| row and column sums of the first pass are summed from
// the same operand values, each lane over its k = lk mod 4, then over the
// four lanes of a row.  P and the reduction are laid out as for
// lx_symv_blk_kernel with NR = 16.
template <bool ABS>
__global__ void __launch_bounds__(256) lx_symv_mm_kernel(const double *__restrict__ A, int64_t lda, int n,
                                                         const double *__restrict__ X, int64_t ldx, int nrhs,
                                                         double *__restrict__ P)
{
    constexpr int NR = 16, NC = NR + (ABS ? 1 : 0), SB = LX_SY_SB, SN = SB * LX_SY_BS, XS = 66;
    typedef lx_mfma<double> MF;
    __shared__ double sA[64][65];  // sA[c][r] = A[I*64 + r, J*64 + c]
    __shared__ double sxI[SB][NR][XS];
    __shared__ double sxJ[SB][NR][XS];
    __shared__ double sred[NC][65];
    const int nb = (n + 63) / 64, ns = (nb + SB - 1) / SB;
    const int64_t nsb = (int64_t)ns * (ns + 1) / 2;
    const int tid = threadIdx.x, l = tid & 63, q = tid >> 6, lr = l & 15, lk = l >> 4;
    for (int64_t sbk = blockIdx.x; sbk < nsb; sbk += gridDim.x) {
        int SI = (int)((sqrt(8.0 * (double)sbk + 1.0) - 1.0) * 0.5);
        while ((int64_t)SI * (SI + 1) / 2 > sbk) --SI;
        while ((int64_t)(SI + 1) * (SI + 2) / 2 <= sbk) ++SI;
        const int SJ = (int)(sbk - (int64_t)SI * (SI + 1) / 2);
        __syncthreads();  // the previous super-block is done with the LDS
        for (int e = tid; e < SB * NR * 64; e += 256) {
            const int k = e % 64, c = (e / 64) % NR, b = e / (64 * NR);
            const int ri = SI * SN + b * 64 + k, rj = SJ * SN + b * 64 + k;
            sxI[b][c][k] = (c < nrhs && ri < n) ? X[(int64_t)c * ldx + ri] : 0.0;
            sxJ[b][c][k] = (c < nrhs && rj < n) ? X[(int64_t)c * ldx + rj] : 0.0;
        }
        typename MF::acc accR[SB], accC[SB];
        double absR[SB], absC[SB];
#pragma unroll
        for (int b = 0; b < SB; ++b) {
            absR[b] = absC[b] = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) accR[b][i] = accC[b][i] = 0.0;
        }
        auto blk_ok = [&](int t) {
            const int I = SI * SB + t / SB, J = SJ * SB + t % SB;
            return J <= I && I < nb;
        };
        double nxt[16];
        auto gload = [&](int t) {
            const int I = SI * SB + t / SB, J = SJ * SB + t % SB;
            const int r0 = I * 64, c0 = J * 64, hr = min(64, n - r0), hc = min(64, n - c0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int e = tid + 256 * i, r = e & 63, c = e >> 6;
                nxt[i] = (r < hr && c < hc && (I != J || r >= c)) ? A[(int64_t)(c0 + c) * lda + r0 + r] : 0.0;
            }
        };
        {
            int t = 0;
            while (t < SB * SB && !blk_ok(t)) ++t;  // uniform
            if (t < SB * SB) gload(t);
        }
#pragma unroll
        for (int bi = 0; bi < SB; ++bi) {
#pragma unroll
            for (int bj = 0; bj < SB; ++bj) {
                const int I = SI * SB + bi, J = SJ * SB + bj;
                if (J > I || I >= nb) continue;  // uniform
                __syncthreads();  // sA is free; also orders the sx writes above
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int e = tid + 256 * i;
                    sA[e >> 6][e & 63] = nxt[i];
                }
                __syncthreads();
                {
                    int t = bi * SB + bj + 1;
                    while (t < SB * SB && !blk_ok(t)) ++t;  // uniform
                    if (t < SB * SB) gload(t);
                }
                const bool diag = I == J;
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) {
                    const int k = 4 * kk + lk;
                    // row part: A[16q + lr][k] x_J[k][lr]
                    const double ar = sA[k][16 * q + lr];
                    accR[bi] = MF::mac(ar, sxJ[bj][lr][k], accR[bi]);
                    // column part: A^T[16q + lr][k] = A[k][16q + lr], strictly lower on the diagonal
                    const double at = (!diag || k > 16 * q + lr) ? sA[16 * q + lr][k] : 0.0;
                    accC[bj] = MF::mac(at, sxI[bi][lr][k], accC[bj]);
                    if (ABS) {  // |A| sums of row / column 16q + lr over the lane's k = lk mod 4
                        absR[bi] += fabs(ar);
                        absC[bj] += fabs(at);
                    }
                }
            }
        }
        // out to P: the MFMA tiles hold rows 16q + lk + 4i of column lr; the |A|
        // sums of row 16q + lr meet over the four lanes lr + 16 lk
        double *Prow = P + ((size_t)SI * ns + SJ) * SN * NC;
        double *Pcol = P + ((size_t)SJ * ns + SI) * SN * NC;
        auto lanes4 = [](double v) {
            v += __shfl_xor(v, 16);
            return v + __shfl_xor(v, 32);
        };
        for (int b = 0; b < SB; ++b) {
            const double ar = ABS ? lanes4(absR[b]) : 0.0, ac = ABS ? lanes4(absC[b]) : 0.0;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; ++i) sred[lr][16 * q + MF::row(lk, i)] = accR[b][i];
            if (ABS && lk == 0) sred[NR][16 * q + lr] = ar;
            __syncthreads();
            if (SI == SJ) {  // uniform: both parts land in P[SI][SI]
#pragma unroll
                for (int i = 0; i < 4; ++i) sred[lr][16 * q + MF::row(lk, i)] += accC[b][i];
                if (ABS && lk == 0) sred[NR][16 * q + lr] += ac;
                __syncthreads();
            }
            for (int e = tid; e < NC * 64; e += 256) {
                const int r = e % 64, c = e / 64;
                Prow[(size_t)c * SN + b * 64 + r] = sred[c][r];
            }
            if (SI != SJ) {
                __syncthreads();
#pragma unroll
                for (int i = 0; i < 4; ++i) sred[lr][16 * q + MF::row(lk, i)] = accC[b][i];
                if (ABS && lk == 0) sred[NR][16 * q + lr] = ac;
                __syncthreads();
                for (int e = tid; e < NC * 64; e += 256) {
                    const int r = e % 64, c = e / 64;
                    Pcol[(size_t)c * SN + b * 64 + r] = sred[c][r];
                }
            }
        }
    }
}

// R = B - y with y_S = sum over SJ = 0.. of P[S][SJ]; rowabs (when non-null)
// = the |A| row sums (column NR of P).  Thread = one row of one right-hand side.
template <int NR, bool ABS>
__global__ void __launch_bounds__(256) lx_symv_reduce_kernel(int n, int nrhs, const double *__restrict__ P,
                                                             const double *__restrict__ B, int64_t ldb,
                                                             double *__restrict__ R, double *__restrict__ rowabs)
{
    constexpr int NC = NR + (ABS ? 1 : 0), SN = LX_SY_SB * LX_SY_BS;
    const int nb = (n + 63) / 64, ns = (nb + LX_SY_SB - 1) / LX_SY_SB;
    const int64_t tot = (int64_t)ns * SN * NC;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
        const int r = (int)(e % SN), c = (int)((e / SN) % NC), S = (int)(e / ((int64_t)SN * NC));
        const int row = S * SN + r;
        if (row >= n) continue;
        if (c < NR && c >= nrhs) continue;
        const double *p = P + (size_t)S * ns * SN * NC + (size_t)c * SN + r;
        double y = 0.0;
        for (int sj = 0; sj < ns; ++sj) y += p[(size_t)sj * SN * NC];
        if (c < NR) R[(int64_t)c * n + row] = B[(int64_t)c * ldb + row] - y;
        else if (rowabs) rowabs[row] = y;
    }
}

// NaN-propagating maximum of one value per thread over a 1024-thread
// workgroup (wave shuffles, then the 16 wave results); the result is valid in
// thread 0.
__device__ __forceinline__ double lx_nanmax(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double lx_wg1024_nanmax(double m)
{
    __shared__ double red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = lx_nanmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = red[0];
        for (int w = 1; w < 16; ++w) m = lx_nanmax(m, red[w]);
    }
    return m;
}

// out[c] = max_i |V[i, c]| (NaN-propagating); grid = ncols, one 1024-thread
// workgroup per column with eight loads in flight per thread
// out[c] = max_i |D_ii V_ic| (D = I without Deq); one workgroup per column
__global__ void __launch_bounds__(1024) lx_colmax_kernel(const double *__restrict__ V, int64_t ldv, int n,
                                                         const double *__restrict__ Deq, double *__restrict__ out)
{
    const double *v = V + (int64_t)blockIdx.x * ldv;
    double m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = 0.0;
    for (int i0 = threadIdx.x; i0 < n; i0 += 8 * 1024) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * 1024;
            if (i < n) m[u] = lx_nanmax(fabs(Deq ? v[i] * Deq[i] : v[i]), m[u]);
        }
    }
#pragma unroll
    for (int u = 1; u < 8; ++u) m[0] = lx_nanmax(m[u], m[0]);
    m[0] = lx_wg1024_nanmax(m[0]);
    if (threadIdx.x == 0) out[blockIdx.x] = m[0];
}

// berr[c] = rmax[c] / (anorm * xmax[c])  (the FP64 convergence test); a zero
// residual is an exact solution (b_c = 0 gives x_c = 0 and r_c = 0), backward
// error 0 rather than 0 / 0
__global__ void lx_berr_kernel(int nrhs, const double *__restrict__ rmax, const double *__restrict__ xmax,
                               const double *__restrict__ anorm, double *__restrict__ berr)
{
    for (int c = threadIdx.x; c < nrhs; c += blockDim.x) berr[c] = rmax[c] == 0.0 ? 0.0 : rmax[c] / (anorm[0] * xmax[c]);
}

// lower triangle of A (FP64) -> Af (FP32), ld n
// Equilibration by powers of two: Deq[i] = 2^-e, e = ceil(p / 2) for
// a_ii = m 2^p (m in [1/2, 1)), so that Deq[i]^2 a_ii lies in [1/4, 1); 1 where
// a_ii is not a positive finite number (the factor then reports it).  For an
// SPD matrix |a_ij| <= sqrt(a_ii a_jj), so every entry of D A D and of its
// Cholesky factor is below 1 in magnitude.
__global__ void lx_equil_kernel(int n, const double *__restrict__ A, int64_t lda, double *__restrict__ Deq)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = A[(int64_t)i * lda + i];
    double d = 1.0;
    if (a > 0.0 && a <= DBL_MAX) {
        int p = 0;
        (void)frexp(a, &p);
        const int e = p >= 0 ? (p + 1) / 2 : -((-p) / 2);
        d = ldexp(1.0, -e);
    }
    Deq[i] = d;
}

// FP32 copy of the lower triangle, of D A D when Deq is given (exact scaling).
// A thread owns rows i = x + k (gridDim.x blockDim.x) and walks the columns
// j = blockIdx.y + m gridDim.y <= i: for each column the rows of a wave are
// consecutive (coalesced), D_i stays in a register and D_j is one address for
// the whole wave.
__global__ void lx_d2s_lower_kernel(const double *__restrict__ A, int64_t lda, int n, const double *__restrict__ Deq,
                                    float *__restrict__ Af)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (Deq) {
            const double di = Deq[i];
            for (int j = blockIdx.y; j <= i; j += gridDim.y)
                Af[(int64_t)j * n + i] = (float)(A[(int64_t)j * lda + i] * (di * Deq[j]));
        } else {
            for (int j = blockIdx.y; j <= i; j += gridDim.y) Af[(int64_t)j * n + i] = (float)A[(int64_t)j * lda + i];
        }
    }
}

// Sf[:, c] = (float)(D V[:, c] / scale[c])   (D = I without Deq)
__global__ void lx_d2s_scaled_kernel(int n, int nrhs, const double *__restrict__ V, int64_t ldv,
                                     const double *__restrict__ scale, const double *__restrict__ Deq,
                                     float *__restrict__ Sf)
{
    int64_t tot = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / n, i = e % n;
        double s = scale[c];
        const double v = Deq ? V[c * ldv + i] * Deq[i] : V[c * ldv + i];
        Sf[c * n + i] = (float)(s > 0.0 ? v / s : 0.0);
    }
}

// X[:, c] (+)= scale[c] * D Sf[:, c]   (FP64 update, itref.c:1084)
__global__ void lx_update_kernel(int n, int nrhs, const float *__restrict__ Sf, const double *__restrict__ scale,
                                 const double *__restrict__ Deq, double *__restrict__ X, int64_t ldx, int accumulate)
{
    int64_t tot = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / n, i = e % n;
        double d = scale[c] * (double)Sf[c * n + i];
        if (Deq) d *= Deq[i];
        X[c * ldx + i] = accumulate ? X[c * ldx + i] + d : d;
    }
}

__global__ void lx_copy_cols_kernel(int n, int ncols, const double *__restrict__ S, int64_t lds,
                                    double *__restrict__ D, int64_t ldd)
{
    int64_t tot = (int64_t)n * ncols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / n, i = e % n;
        D[c * ldd + i] = S[c * lds + i];
    }
}

// out[0] = max_i v[i] (NaN-propagating, v >= 0); one 1024-thread workgroup
__global__ void __launch_bounds__(1024) lx_max_kernel(int n, const double *__restrict__ v, double *__restrict__ out)
{
    double m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = 0.0;
    for (int i0 = threadIdx.x; i0 < n; i0 += 8 * 1024) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * 1024;
            if (i < n) m[u] = lx_nanmax(v[i], m[u]);
        }
    }
#pragma unroll
    for (int u = 1; u < 8; ++u) m[0] = lx_nanmax(m[u], m[0]);
    m[0] = lx_wg1024_nanmax(m[0]);
    if (threadIdx.x == 0) out[0] = m[0];
}

static inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// r = b - A x for nc <= NR right-hand sides in one pass over A; rowabs (when
// non-null) receives the row sums of |A|.
template <int NR>
static int residual_nr(int n, int nc, const double *A, int64_t lda, const double *X, int64_t ldx, const double *B,
                       int64_t ldb, double *R, double *rowabs, double *P, hipStream_t s)
{
    const int nb = (n + 63) / 64, ns = (nb + LX_SY_SB - 1) / LX_SY_SB;
    const int64_t nsb = (int64_t)ns * (ns + 1) / 2;
    const int grid = (int)std::min<int64_t>(nsb, 2048);
    const int64_t red = (int64_t)ns * LX_SY_SB * LX_SY_BS * (NR + 1);
    if constexpr (NR == 16) {  // the matrix-core pass
        if (rowabs) {
            hipLaunchKernelGGL((lx_symv_mm_kernel<true>), dim3(grid), dim3(256), 0, s, A, lda, n, X, ldx, nc, P);
            hipLaunchKernelGGL((lx_symv_reduce_kernel<NR, true>), dim3(grid_for(red)), dim3(256), 0, s, n, nc, P, B,
                               ldb, R, rowabs);
        } else {
            hipLaunchKernelGGL((lx_symv_mm_kernel<false>), dim3(grid), dim3(256), 0, s, A, lda, n, X, ldx, nc, P);
            hipLaunchKernelGGL((lx_symv_reduce_kernel<NR, false>), dim3(grid_for(red)), dim3(256), 0, s, n, nc, P, B,
                               ldb, R, nullptr);
        }
        return 0;
    }
    if (rowabs) {
        hipLaunchKernelGGL((lx_symv_blk_kernel<NR, true>), dim3(grid), dim3(256), 0, s, A, lda, n, X, ldx, nc, P);
        hipLaunchKernelGGL((lx_symv_reduce_kernel<NR, true>), dim3(grid_for(red)), dim3(256), 0, s, n, nc, P, B, ldb, R,
                           rowabs);
    } else {
        hipLaunchKernelGGL((lx_symv_blk_kernel<NR, false>), dim3(grid), dim3(256), 0, s, A, lda, n, X, ldx, nc, P);
        hipLaunchKernelGGL((lx_symv_reduce_kernel<NR, false>), dim3(grid_for(red)), dim3(256), 0, s, n, nc, P, B, ldb,
                           R, nullptr);
    }
    return 0;
}

static int residual(int n, int nrhs, const double *A, int64_t lda, const double *X, int64_t ldx, const double *B,
                    int64_t ldb, double *R, double *rowabs, hipStream_t s)
{
    const int nb = (n + 63) / 64, ns = (nb + LX_SY_SB - 1) / LX_SY_SB;
    // one column on the vector units; 2-4 four at a time; more in passes of 16
    // on the matrix cores
    const int nrp = nrhs == 1 ? 1 : nrhs <= 4 ? 4 : 16;
    double *P = (double *)lx_scratch(8, sizeof(double) * (size_t)ns * ns * LX_SY_SB * LX_SY_BS * (nrp + 1));
    if (!P) return LINOMPSS_ERR_NOMEM;
    for (int c0 = 0; c0 < nrhs; c0 += nrp) {
        const int nc = std::min(nrp, nrhs - c0);
        double *ra = c0 == 0 ? rowabs : nullptr;
        if (nrp == 1)
            residual_nr<1>(n, nc, A, lda, X + c0 * ldx, ldx, B + c0 * ldb, ldb, R + (int64_t)c0 * n, ra, P, s);
        else if (nrp == 4)
            residual_nr<4>(n, nc, A, lda, X + c0 * ldx, ldx, B + c0 * ldb, ldb, R + (int64_t)c0 * n, ra, P, s);
        else
            residual_nr<16>(n, nc, A, lda, X + c0 * ldx, ldx, B + c0 * ldb, ldb, R + (int64_t)c0 * n, ra, P, s);
    }
    LX_HIP_TRY(hipGetLastError());
    return 0;
}

static std::mutex g_ir_mu;

extern "C" int linompss_dposv_ir_dev(int n, int nrhs, int b, const double *dA, int64_t lda, const double *dB,
                                     int64_t ldb, double *dX, int64_t ldx, void *stream, const linompss_ir_opts *opts,
                                     linompss_ir_status *st)
{
    if (n < 0) return -1;
    if (nrhs < 0) return -2;
    if (b <= 0) return -3;
    if (lda < std::max(1, n)) return -5;
    if (ldb < std::max(1, n)) return -7;
    if (ldx < std::max(1, n)) return -9;
    linompss_ir_status local = {};
    if (!st) st = &local;
    *st = linompss_ir_status{};
    if (n == 0 || nrhs == 0) return 0;
    linompss_ir_opts o = {30, 1e-14, 0.5, 0};
    if (opts) o = *opts;
    if (o.update_prec < 0 || o.update_prec > 2) return -11;  // 0 FP32, 1 split-BF16, 2 split-FP16 updates
    std::lock_guard<std::mutex> g(g_ir_mu);
    hipStream_t s = (hipStream_t)stream;
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
    // per-column values: [0] anorm, then rmax, xmax, berr, bmax, smax (nrhs each)
    const size_t nsmall = (1 + 5 * (size_t)nrhs + 63) / 64 * 64;
    float *Af = (float *)lx_scratch(4, sizeof(float) * nn);
    char *ws = (char *)lx_scratch(5, sizeof(double) * (nr + 2 * (size_t)n + nsmall) + sizeof(float) * nr + 1024);
    if (!Af || !ws) return LINOMPSS_ERR_NOMEM;
    double *R = (double *)ws;
    double *rowabs = R + nr;
    // the FP16 split needs |L| <= 16: factor D A D (unit-range diagonal)
    double *Deq = o.update_prec == 2 ? rowabs + n : nullptr;
    double *small = rowabs + 2 * (size_t)n;
    float *Sf = (float *)(small + nsmall);
    double *anorm = small, *rmax = small + 1, *xmax = rmax + nrhs, *berr = xmax + nrhs, *bmax = berr + nrhs;
    double *smax = bmax + nrhs;  // max |D r| per column: the scale of the equilibrated corrections
    const size_t nhost = 1 + 3 * (size_t)nrhs;  // anorm, rmax, xmax, berr
    std::vector<double> h_small(nhost);
    // 1) FP32 copy of the lower triangle, FP32 factor
    {
        dim3 grid(std::max(1, std::min(64, (n + 255) / 256)), std::min(n, 4096));
        if (Deq) hipLaunchKernelGGL(lx_equil_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, dA, lda, Deq);
        hipLaunchKernelGGL(lx_d2s_lower_kernel, grid, dim3(256), 0, s, dA, lda, n, Deq, Af);
        LX_HIP_TRY(hipGetLastError());
    }
    linompss_exec_opts eo = {1, 128, 128, 0, 30.0, 0, o.update_prec};
    int info = 0;
    int rc = linompss_chol_dev(1, 0, n, b, Af, n, s, &eo);
    linompss_exec_stats es;
    linompss_last_exec_stats(&es);
    st->t_factor_s = es.exec_ms * 1e-3;
    if (rc < 0) return rc;
    info = rc;
    bool fallback = info != 0;
    double prev = INFINITY, cur_berr = INFINITY;
    // the 128-row sweeps, their step matrices computed once from this factor
    // and kept for every solve of the loop
    float *prep = nullptr;
    if (!fallback && lx_sw2_enabled(1, n, nrhs)) {
        std::lock_guard<std::mutex> g(lx_dev_mutex());
        prep = (float *)lx_scratch(14, lx_sw2_prep_bytes(n));
        if (!prep) return LINOMPSS_ERR_NOMEM;
        rc = lx_sw2_prep(n, Af, n, prep, s);
        if (rc) return rc;
    }
    auto fsolve = [&]() -> int {
        if (!prep) return linompss_potrs_dev(1, n, nrhs, b, Af, n, Sf, n, s, &eo);
        std::lock_guard<std::mutex> g(lx_dev_mutex());
        return lx_sw2_sweeps(n, nrhs, Af, n, Sf, n, 3, prep, s);
    };
    if (!fallback) {
        // x0 = solve(b) with per-column scaling by ||b_c||_inf
        // the FP32 right-hand side D b / max|D b| lies in [-1, 1] whatever D is
        hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, dB, ldb, n, Deq, bmax);
        hipLaunchKernelGGL(lx_d2s_scaled_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, dB, ldb, bmax,
                           Deq, Sf);
        rc = fsolve();
        if (rc) return rc;
        hipLaunchKernelGGL(lx_update_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, Sf, bmax, Deq, dX,
                           ldx, 0);
        for (int it = 0;; ++it) {
            rc = residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, it == 0 ? rowabs : nullptr, s);
            if (rc) return rc;
            if (it == 0) hipLaunchKernelGGL(lx_max_kernel, dim3(1), dim3(1024), 0, s, n, rowabs, anorm);
            hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, R, (int64_t)n, n, nullptr, rmax);
            hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, dX, ldx, n, nullptr, xmax);
            hipLaunchKernelGGL(lx_berr_kernel, dim3(1), dim3(256), 0, s, nrhs, rmax, xmax, anorm, berr);
            if (Deq) hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, R, (int64_t)n, n, Deq, smax);
            LX_HIP_TRY(hipGetLastError());
            LX_HIP_TRY(hipMemcpyAsync(h_small.data(), small, sizeof(double) * nhost, hipMemcpyDeviceToHost, s));
            LX_HIP_TRY(hipStreamSynchronize(s));
            double worst = 0.0, rn = 0.0;
            for (int c = 0; c < nrhs; ++c) {
                double e = h_small[1 + 2 * (size_t)nrhs + c], r = h_small[1 + c];
                if (!(e <= worst)) worst = e;
                if (!(r <= rn)) rn = r;
            }
            cur_berr = worst;
            if (cur_berr <= o.tol) break;
            if (it >= o.max_sweeps || !(rn <= o.stall_ratio * prev) || std::isnan(rn)) {
                fallback = true;
                break;
            }
            prev = rn;
            // d = solve(r / ||r||); x += ||r|| d
            hipLaunchKernelGGL(lx_d2s_scaled_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, R,
                               (int64_t)n, Deq ? smax : rmax, Deq, Sf);
            rc = fsolve();
            if (rc) return rc;
            hipLaunchKernelGGL(lx_update_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, Sf,
                               Deq ? smax : rmax, Deq, dX, ldx, 1);
            st->sweeps = it + 1;
        }
    }
    if (fallback) {
        double *Ad = (double *)lx_scratch(6, sizeof(double) * nn);
        if (!Ad) return LINOMPSS_ERR_NOMEM;
        hipLaunchKernelGGL(lx_copy_cols_kernel, dim3(grid_for((int64_t)nn)), dim3(256), 0, s, n, n, dA, lda, Ad,
                           (int64_t)n);
        hipLaunchKernelGGL(lx_copy_cols_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, dB, ldb, dX,
                           ldx);
        LX_HIP_TRY(hipGetLastError());
        rc = linompss_posv_dev(0, n, nrhs, b, Ad, n, dX, ldx, s, &eo);
        if (rc < 0) return rc;
        info = rc;
        rc = residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, rowabs, s);
        if (rc) return rc;
        hipLaunchKernelGGL(lx_max_kernel, dim3(1), dim3(1024), 0, s, n, rowabs, anorm);
        hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, R, (int64_t)n, n, nullptr, rmax);
        hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, dX, ldx, n, nullptr, xmax);
        hipLaunchKernelGGL(lx_berr_kernel, dim3(1), dim3(256), 0, s, nrhs, rmax, xmax, anorm, berr);
        LX_HIP_TRY(hipMemcpyAsync(h_small.data(), small, sizeof(double) * nhost, hipMemcpyDeviceToHost, s));
        LX_HIP_TRY(hipStreamSynchronize(s));
        cur_berr = 0.0;
        for (int c = 0; c < nrhs; ++c)
            if (!(h_small[1 + 2 * (size_t)nrhs + c] <= cur_berr)) cur_berr = h_small[1 + 2 * (size_t)nrhs + c];
    }
    st->fallback = fallback ? 1 : 0;
    st->info = info;
    st->berr = cur_berr;
    st->t_total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return info;
}

extern "C" int linompss_dposv_ir(int n, int nrhs, int b, const double *A, int lda, const double *B, int ldb, double *X,
                                 int ldx, const linompss_ir_opts *opts, linompss_ir_status *st)
{
    if (n < 0) return -1;
    if (nrhs < 0) return -2;
    if (b <= 0) return -3;
    if (lda < std::max(1, n)) return -5;
    if (ldb < std::max(1, n)) return -7;
    if (ldx < std::max(1, n)) return -9;
    if (n == 0 || nrhs == 0) return 0;
    double *dA = nullptr, *dB = nullptr, *dX = nullptr;
    const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
    if (hipMalloc(&dA, sizeof(double) * nn) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    if (hipMalloc(&dB, sizeof(double) * nr) != hipSuccess || hipMalloc(&dX, sizeof(double) * nr) != hipSuccess) {
        (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
        return LINOMPSS_ERR_NOMEM;
    }
    int rc = 0;
    if (hipMemcpy2D(dA, sizeof(double) * n, A, sizeof(double) * (size_t)lda, sizeof(double) * n, n,
                    hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy2D(dB, sizeof(double) * n, B, sizeof(double) * (size_t)ldb, sizeof(double) * n, nrhs,
                    hipMemcpyHostToDevice) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc) rc = linompss_dposv_ir_dev(n, nrhs, b, dA, n, dB, n, dX, n, nullptr, opts, st);
    if (rc >= 0 && hipMemcpy2D(X, sizeof(double) * (size_t)ldx, dX, sizeof(double) * n, sizeof(double) * n, nrhs,
                               hipMemcpyDeviceToHost) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dX);
    return rc;
}

// Normwise backward error max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf)
// in FP64 on the device, A symmetric with its lower triangle stored.
extern "C" int linompss_dberr_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dX, int64_t ldx,
                                  const double *dB, int64_t ldb, double *berr_out)
{
    if (n <= 0) return -1;
    if (nrhs <= 0) return -2;
    if (lda < n) return -4;
    if (ldx < n) return -6;
    if (ldb < n) return -8;
    std::lock_guard<std::mutex> g(g_ir_mu);
    const size_t nr = (size_t)n * nrhs;
    char *ws = (char *)lx_scratch(7, sizeof(double) * (nr + (size_t)n + 1 + 3 * (size_t)nrhs + 64));
    if (!ws) return LINOMPSS_ERR_NOMEM;
    double *R = (double *)ws, *rowabs = R + nr;
    double *small = rowabs + n;
    double *anorm = small, *rmax = small + 1, *xmax = rmax + nrhs, *berr = xmax + nrhs;
    hipStream_t s = nullptr;
    int rc = residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, rowabs, s);
    if (rc) return rc;
    hipLaunchKernelGGL(lx_max_kernel, dim3(1), dim3(1024), 0, s, n, rowabs, anorm);
    hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, R, (int64_t)n, n, nullptr, rmax);
    hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(1024), 0, s, dX, ldx, n, nullptr, xmax);
    hipLaunchKernelGGL(lx_berr_kernel, dim3(1), dim3(256), 0, s, nrhs, rmax, xmax, anorm, berr);
    std::vector<double> h(nrhs);
    LX_HIP_TRY(hipMemcpy(h.data(), berr, sizeof(double) * nrhs, hipMemcpyDeviceToHost));
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c)
        if (!(h[c] <= worst)) worst = h[c];
    *berr_out = worst;
    return 0;
}
