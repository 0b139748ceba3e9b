// lx_gp.hip -- the squared-exponential Gaussian-process kernel matrix and the
// contraction of its adjoint with the kernel's parameter derivatives, on the
// device.  With these the GP likelihood's forward and backward passes keep K
// on the GPU: only the points, y and scalars cross PCIe.
//
// The reference has no Gaussian-process code; these feed its factor
// (task_dpotrf, tasks_potrf.c:20-24, through CHOL_RL) and the solve adjoints.
//
//   K_ij = sf2 exp(-|x_i - x_j|^2 / (2 l^2)) + noise delta_ij
//
// Given Kbar in the lower representation (dLoss = sum_{i>=j} Kbar_ij dK_ij):
//   dl     = sum_{i>j} Kbar_ij k_ij |x_i - x_j|^2 / l^3      (k = sf2 exp(...))
//   dsf2   = sum_{i>j} Kbar_ij k_ij / sf2 + sum_i Kbar_ii
//   dnoise = sum_i Kbar_ii
//   dX_i   = -sum_{j != i} W_ij (x_i - x_j),  W_ij = W_ji = Kbar_ij k_ij / l^2 (i > j)
// Sums run over 64 x 64 tiles of the lower triangle, each tile's partials in
// its own slot, then a fixed-order reduction: the results do not depend on
// scheduling.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "linompss.h"
#include "lx_device.hpp"
#include "lx_internal.hpp"

#define LX_GP_DMAX 16  // point dimension handled per launch

// One 64 x 64 tile per workgroup: rows of tile (bi, bj), both triangles
// written (tile (bi, bj) and its mirror), bj <= bi.
__global__ void __launch_bounds__(256) lx_gp_kernel_k(int n, int d, const double *__restrict__ X, int64_t ldx,
                                                      double inv2l2, double sf2, double noise, double *__restrict__ K,
                                                      int64_t ldk)
{
    const int t = blockIdx.x;  // lower tile index: bi (bi + 1) / 2 + bj
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    __shared__ double xr[LX_GP_DMAX][64], xc[LX_GP_DMAX][64];
    const int tid = threadIdx.x;
    for (int e = tid; e < d * 64; e += 256) {
        const int q = e / 64, r = e % 64;
        const int gi = min(bi * 64 + r, n - 1), gj = min(bj * 64 + r, n - 1);
        xr[q][r] = X[(int64_t)q * ldx + gi];
        xc[q][r] = X[(int64_t)q * ldx + gj];
    }
    __syncthreads();
    const int r = tid & 63;
    const int gi = bi * 64 + r;
    for (int c = tid >> 6; c < 64; c += 4) {
        const int gj = bj * 64 + c;
        if (gi >= n || gj >= n) continue;
        double s = 0.0;
        for (int q = 0; q < d; ++q) {
            const double df = xr[q][r] - xc[q][c];
            s += df * df;
        }
        const double v = sf2 * exp(-s * inv2l2) + (gi == gj ? noise : 0.0);
        K[(int64_t)gj * ldk + gi] = v;
        if (bi != bj) K[(int64_t)gi * ldk + gj] = v;  // the mirror tile (the diagonal tile writes both itself)
    }
}

// Partials of one lower tile: part[t * (2 + 128 d) + ...] =
//   [0] sum Kbar k r^2 (i > j), [1] sum Kbar k (i > j, / sf2 later),
//   [2 + q * 64 + r]: row part of dX (rows of bi), [2 + 64 d + q * 64 + c]: column part (rows of bj).
// The diagonal entries' sums (dnoise, the sf2 diagonal term) come from a separate pass.
__global__ void __launch_bounds__(256) lx_gp_grad_part(int n, int d, const double *__restrict__ X, int64_t ldx,
                                                       double inv2l2, double sf2, const double *__restrict__ Kb,
                                                       int64_t ldkb, double *__restrict__ part)
{
    const int t = blockIdx.x;
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    __shared__ double xr[LX_GP_DMAX][64], xc[LX_GP_DMAX][64];
    __shared__ double w[64][65];  // W_ij = Kbar_ij k_ij / l^2 ... scaled below
    __shared__ double red[2][256];
    const int tid = threadIdx.x;
    for (int e = tid; e < d * 64; e += 256) {
        const int q = e / 64, r = e % 64;
        const int gi = min(bi * 64 + r, n - 1), gj = min(bj * 64 + r, n - 1);
        xr[q][r] = X[(int64_t)q * ldx + gi];
        xc[q][r] = X[(int64_t)q * ldx + gj];
    }
    __syncthreads();
    const int r = tid & 63;
    const int gi = bi * 64 + r;
    double s_r2 = 0.0, s_k = 0.0;
#pragma unroll 1
    for (int c = tid >> 6; c < 64; c += 4) {
        const int gj = bj * 64 + c;
        double wv = 0.0;
        if (gi < n && gj < n && gi > gj) {
            double s = 0.0;
            for (int q = 0; q < d; ++q) {
                const double df = xr[q][r] - xc[q][c];
                s += df * df;
            }
            const double kb = Kb[(int64_t)gj * ldkb + gi] * sf2 * exp(-s * inv2l2);
            s_r2 += kb * s;
            s_k += kb;
            wv = kb;
        }
        w[r][c] = wv;
    }
    red[0][tid] = s_r2;
    red[1][tid] = s_k;
    __syncthreads();
    double *P = part + (int64_t)t * (2 + 128 * d);
    if (tid < 2) {
        double a = 0.0;
#pragma unroll 8
        for (int q = 0; q < 256; ++q) a += red[tid][q];  // fixed order
        P[tid] = a;
    }
    // dX partials (times 1 / l^2 later): row i of bi: -sum_c w[i][c] (x_i - x_c);
    // row c of bj: +sum_i w[i][c] (x_i - x_c)
    for (int e = tid; e < 64 * d; e += 256) {
        const int q = e / 64, m = e % 64;
        double a = 0.0, b = 0.0;
#pragma unroll 4
        for (int c = 0; c < 64; ++c) a -= w[m][c] * (xr[q][m] - xc[q][c]);
#pragma unroll 4
        for (int i = 0; i < 64; ++i) b += w[i][m] * (xr[q][i] - xc[q][m]);
        P[2 + q * 64 + m] = a;
        P[2 + 64 * d + q * 64 + m] = b;
    }
}

// Fixed-order reduction of the partials: out[0] = sum r^2 term, out[1] = sum k
// term, out[2] = trace of Kbar; dX (n x d, optional) from the row and column parts.
__global__ void __launch_bounds__(256) lx_gp_grad_reduce(int n, int d, int nb, const double *__restrict__ part,
                                                         const double *__restrict__ Kb, int64_t ldkb, double l2inv,
                                                         double *__restrict__ out, double *__restrict__ dX,
                                                         int64_t lddx)
{
    const int tid = threadIdx.x;
    if (blockIdx.x == 0) {
        __shared__ double red[3][256];
        const int nt = nb * (nb + 1) / 2;
        double a = 0.0, b = 0.0, tr = 0.0;
        for (int t = tid; t < nt; t += 256) {
            a += part[(int64_t)t * (2 + 128 * d)];
            b += part[(int64_t)t * (2 + 128 * d) + 1];
        }
        for (int i = tid; i < n; i += 256) tr += Kb[(int64_t)i * ldkb + i];
        red[0][tid] = a;
        red[1][tid] = b;
        red[2][tid] = tr;
        __syncthreads();
        if (tid < 3) {
            double s = 0.0;
#pragma unroll 8
            for (int q = 0; q < 256; ++q) s += red[tid][q];
            out[tid] = s;
        }
        return;
    }
    if (!dX) return;
    // one thread per (point, coordinate)
    const int64_t e = (int64_t)(blockIdx.x - 1) * 256 + tid;
    if (e >= (int64_t)n * d) return;
    const int i = (int)(e % n), q = (int)(e / n), bi = i / 64, m = i % 64;
    double s = 0.0;
    for (int bj = 0; bj <= bi; ++bj)  // tiles (bi, bj): i is a row
        s += part[(int64_t)(bi * (bi + 1) / 2 + bj) * (2 + 128 * d) + 2 + q * 64 + m];
    for (int bk = bi; bk < nb; ++bk)  // tiles (bk, bi): i is a column
        s += part[(int64_t)(bk * (bk + 1) / 2 + bi) * (2 + 128 * d) + 2 + 64 * d + q * 64 + m];
    dX[(int64_t)q * lddx + i] = s * l2inv;
}

namespace {
std::mutex lx_gp_mu;  // the partials' scratch slot
}

extern "C" int linompss_gp_kernel_dev(int n, int d, const double *dX, int64_t ldx, double ell, double sf2,
                                      double noise, double *dK, int64_t ldk, void *stream)
{
    if (n < 0) return -1;
    if (d < 1 || d > LX_GP_DMAX) return -2;
    if (!dX && n > 0) return -3;
    if (ldx < std::max(1, n)) return -4;
    if (!(ell > 0.0)) return -5;
    if (ldk < std::max(1, n)) return -9;
    if (n == 0) return 0;
    const int nb = (n + 63) / 64;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lx_gp_kernel_k, dim3(nb * (nb + 1) / 2), dim3(256), 0, s, n, d, dX, ldx,
                       1.0 / (2.0 * ell * ell), sf2, noise, dK, ldk);
    LX_HIP_TRY(hipGetLastError());
    LX_HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

extern "C" int linompss_gp_grad_dev(int n, int d, const double *dX, int64_t ldx, double ell, double sf2,
                                    const double *dKbar, int64_t ldkb, double *grads, double *dXbar, int64_t ldxb,
                                    void *stream)
{
    if (n < 0) return -1;
    if (d < 1 || d > LX_GP_DMAX) return -2;
    if (!dX && n > 0) return -3;
    if (ldx < std::max(1, n)) return -4;
    if (!(ell > 0.0)) return -5;
    if (!(sf2 > 0.0)) return -6;
    if (!dKbar && n > 0) return -7;
    if (ldkb < std::max(1, n)) return -8;
    if (!grads) return -9;
    if (dXbar && ldxb < std::max(1, n)) return -11;
    grads[0] = grads[1] = grads[2] = 0.0;
    if (n == 0) return 0;
    const int nb = (n + 63) / 64, nt = nb * (nb + 1) / 2;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(lx_gp_mu);
    double *part = (double *)lx_scratch(12, sizeof(double) * ((size_t)nt * (2 + 128 * d) + 4));
    if (!part) return LINOMPSS_ERR_NOMEM;
    double *out = part + (size_t)nt * (2 + 128 * d);
    hipLaunchKernelGGL(lx_gp_grad_part, dim3(nt), dim3(256), 0, s, n, d, dX, ldx, 1.0 / (2.0 * ell * ell), sf2,
                       dKbar, ldkb, part);
    const int nred = 1 + (dXbar ? (int)(((int64_t)n * d + 255) / 256) : 0);
    hipLaunchKernelGGL(lx_gp_grad_reduce, dim3(nred), dim3(256), 0, s, n, d, nb, part, dKbar, ldkb,
                       1.0 / (ell * ell), out, dXbar, ldxb);
    LX_HIP_TRY(hipGetLastError());
    double h[3];
    LX_HIP_TRY(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, s));
    LX_HIP_TRY(hipStreamSynchronize(s));
    grads[0] = h[0] / (ell * ell * ell);  // dl
    grads[1] = h[1] / sf2 + h[2];         // dsf2
    grads[2] = h[2];                      // dnoise
    return 0;
}
