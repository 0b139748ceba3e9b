// lx_ir.hip -- mixed-precision iterative refinement on the device.
//
// Loop contract of the reference's itref (src/itref/itref.c:999-1101):
//   x0 = solve(b); r = b - A x (FP64, itref.c:1010-1020)
//   for i < refit: test ||r||; d = solve(r) (itref.c:1062); x += d
//                  (bblas_sdcpaxpy, itref.c:1084); r = b - A x (itref.c:1094)
// Here solve() is the FP32 tiled Cholesky factor + sweeps of this library,
// and the FP64 pieces (residual, update, backward-error test) are the kernels
// below.  A is symmetric and only its lower triangle is read: the residual is
// r = b - (tril(A) x + stril(A)^T x), two coalesced passes over the lower
// triangle that also produce ||A||_inf on the first sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <vector>

#include "linompss.h"
#include "lx_internal.hpp"

#define LX_IR_NR 16  // right-hand sides per kernel pass

// y1 = tril(A) X, rowabs = sum_{j<=i} |a_ij|  (rows: lanes; columns split by wave)
__global__ void __launch_bounds__(256)
    lx_symv_rows_kernel(const double *__restrict__ A, int64_t lda, int n, const double *__restrict__ X, int64_t ldx,
                        int nrhs, double *__restrict__ Y1, double *__restrict__ rowabs)
{
    __shared__ double red[4][LX_IR_NR + 1][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * 64, row = r0 + lane;
    const int jend = std::min(n, r0 + 64);  // columns 0 .. jend-1 (masked j <= row)
    const int chunk = (jend + 3) / 4;
    const int jb = w * chunk, je = std::min(jend, jb + chunk);
    double acc[LX_IR_NR];
#pragma unroll
    for (int c = 0; c < LX_IR_NR; ++c) acc[c] = 0.0;
    double aabs = 0.0;
    if (row < n) {
        const double *Ar = A + row;
        int j = jb;
        for (; j < je; ++j) {
            double a = (j <= row) ? Ar[(int64_t)j * lda] : 0.0;
            aabs += fabs(a);
#pragma unroll
            for (int c = 0; c < LX_IR_NR; ++c)
                if (c < nrhs) acc[c] += a * X[(int64_t)c * ldx + j];
        }
    }
#pragma unroll
    for (int c = 0; c < LX_IR_NR; ++c) red[w][c][lane] = acc[c];
    red[w][LX_IR_NR][lane] = aabs;
    __syncthreads();
    if (w == 0 && row < n) {
        for (int c = 0; c < nrhs; ++c)
            Y1[(int64_t)c * n + row] = (red[0][c][lane] + red[1][c][lane]) + (red[2][c][lane] + red[3][c][lane]);
        if (rowabs)
            rowabs[row] = (red[0][LX_IR_NR][lane] + red[1][LX_IR_NR][lane]) +
                          (red[2][LX_IR_NR][lane] + red[3][LX_IR_NR][lane]);
    }
}

// y2_j = sum_{i>j} a_ij x_i, colabs_j = sum_{i>j} |a_ij|  (one wave per column)
__global__ void __launch_bounds__(256)
    lx_symv_cols_kernel(const double *__restrict__ A, int64_t lda, int n, const double *__restrict__ X, int64_t ldx,
                        int nrhs, double *__restrict__ Y2, double *__restrict__ colabs)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int j = blockIdx.x * 4 + w; j < n; j += gridDim.x * 4) {
        double acc[LX_IR_NR];
#pragma unroll
        for (int c = 0; c < LX_IR_NR; ++c) acc[c] = 0.0;
        double aabs = 0.0;
        const double *Aj = A + (int64_t)j * lda;
        for (int i = j + 1 + lane; i < n; i += 64) {
            double a = Aj[i];
            aabs += fabs(a);
#pragma unroll
            for (int c = 0; c < LX_IR_NR; ++c)
                if (c < nrhs) acc[c] += a * X[(int64_t)c * ldx + i];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int c = 0; c < LX_IR_NR; ++c)
                if (c < nrhs) acc[c] += __shfl_xor(acc[c], off);
            aabs += __shfl_xor(aabs, off);
        }
        if (lane == 0) {
            for (int c = 0; c < nrhs; ++c) Y2[(int64_t)c * n + j] = acc[c];
            if (colabs) colabs[j] = aabs;
        }
    }
}

// R = B - (Y1 + Y2); optional rowsum = rowabs + colabs
__global__ void lx_resid_combine_kernel(int n, int nrhs, const double *__restrict__ B, int64_t ldb,
                                        const double *__restrict__ Y1, const double *__restrict__ Y2,
                                        double *__restrict__ R, double *__restrict__ rowabs,
                                        const double *__restrict__ colabs)
{
    int64_t tot = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / n, i = e % n;
        R[c * n + i] = B[c * ldb + i] - (Y1[c * n + i] + Y2[c * n + i]);
        if (rowabs && c == 0) rowabs[i] += colabs[i];
    }
}

// out[c] = max_i |V[i, c]| (NaN-propagating); grid = ncols, one WG per column
__global__ void __launch_bounds__(256) lx_colmax_kernel(const double *__restrict__ V, int64_t ldv, int n,
                                                        double *__restrict__ out)
{
    __shared__ double red[256];
    const double *v = V + (int64_t)blockIdx.x * ldv;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double a = fabs(v[i]);
        m = (a > m || a != a) ? a : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = red[threadIdx.x + s], b = red[threadIdx.x];
            red[threadIdx.x] = (a > b || a != a) ? a : b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// berr[c] = rmax[c] / (anorm * xmax[c])  (the FP64 convergence test)
__global__ void lx_berr_kernel(int nrhs, const double *__restrict__ rmax, const double *__restrict__ xmax,
                               const double *__restrict__ anorm, double *__restrict__ berr)
{
    int c = threadIdx.x;
    if (c < nrhs) berr[c] = rmax[c] / (anorm[0] * xmax[c]);
}

// lower triangle of A (FP64) -> Af (FP32), ld n
__global__ void lx_d2s_lower_kernel(const double *__restrict__ A, int64_t lda, int n, float *__restrict__ Af)
{
    for (int j = blockIdx.y; j < n; j += gridDim.y)
        for (int i = j + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            Af[(int64_t)j * n + i] = (float)A[(int64_t)j * lda + i];
}

// Sf[:, c] = (float)(V[:, c] * inv[c])
__global__ void lx_d2s_scaled_kernel(int n, int nrhs, const double *__restrict__ V, int64_t ldv,
                                     const double *__restrict__ scale, float *__restrict__ Sf)
{
    int64_t tot = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / n, i = e % n;
        double s = scale[c];
        Sf[c * n + i] = (float)(s > 0.0 ? V[c * ldv + i] / s : 0.0);
    }
}

// X[:, c] (+)= scale[c] * Sf[:, c]   (FP64 update, itref.c:1084)
__global__ void lx_update_kernel(int n, int nrhs, const float *__restrict__ Sf, const double *__restrict__ scale,
                                 double *__restrict__ X, int64_t ldx, int accumulate)
{
    int64_t tot = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / n, i = e % n;
        double d = scale[c] * (double)Sf[c * n + i];
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

__global__ void lx_max_kernel(int n, const double *__restrict__ v, double *__restrict__ out)
{
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double a = v[i];
        m = (a > m || a != a) ? a : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = red[threadIdx.x + s], b = red[threadIdx.x];
            red[threadIdx.x] = (a > b || a != a) ? a : b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

static inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// r = b - A x for nrhs <= LX_IR_NR columns at a time; rowsum (when non-null)
// receives ||A||'s row sums.
static int residual(int n, int nrhs, const double *A, int64_t lda, const double *X, int64_t ldx, const double *B,
                    int64_t ldb, double *R, double *Y1, double *Y2, double *rowabs, double *colabs, hipStream_t s)
{
    for (int c0 = 0; c0 < nrhs; c0 += LX_IR_NR) {
        int nc = std::min(LX_IR_NR, nrhs - c0);
        bool nrm = rowabs && c0 == 0;
        hipLaunchKernelGGL(lx_symv_rows_kernel, dim3((n + 63) / 64), dim3(256), 0, s, A, lda, n, X + c0 * ldx, ldx,
                           nc, Y1, nrm ? rowabs : nullptr);
        hipLaunchKernelGGL(lx_symv_cols_kernel, dim3(std::min(8192, (n + 3) / 4)), dim3(256), 0, s, A, lda, n,
                           X + c0 * ldx, ldx, nc, Y2, nrm ? colabs : nullptr);
        hipLaunchKernelGGL(lx_resid_combine_kernel, dim3(grid_for((int64_t)n * nc)), dim3(256), 0, s, n, nc,
                           B + c0 * ldb, ldb, Y1, Y2, R + (int64_t)c0 * n, nrm ? rowabs : nullptr, colabs);
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
    if (o.update_prec != 0) return -11;  // low-precision trailing updates are not built yet
    std::lock_guard<std::mutex> g(g_ir_mu);
    hipStream_t s = (hipStream_t)stream;
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
    float *Af = (float *)lx_scratch(4, sizeof(float) * nn);
    char *ws = (char *)lx_scratch(5, sizeof(double) * (3 * nr + 2 * (size_t)n + 64) + sizeof(float) * nr + 1024);
    if (!Af || !ws) return LINOMPSS_ERR_NOMEM;
    double *R = (double *)ws;
    double *Y1 = R + nr, *Y2 = Y1 + nr;
    double *rowabs = Y2 + nr, *colabs = rowabs + n;
    double *small = colabs + n;  // [0] anorm, [1..16] rmax, [17..32] xmax, [33..48] berr, [49..] bmax
    float *Sf = (float *)(small + 128);
    double *anorm = small, *rmax = small + 1, *xmax = small + 1 + 16, *berr = small + 1 + 32, *bmax = small + 1 + 48;
    if (nrhs > 16) {
        // small per-column arrays sized for 16 columns
        return -2;
    }
    // 1) FP32 copy of the lower triangle, FP32 factor
    {
        dim3 grid(std::max(1, std::min(64, (n + 255) / 256)), std::min(n, 4096));
        hipLaunchKernelGGL(lx_d2s_lower_kernel, grid, dim3(256), 0, s, dA, lda, n, Af);
        LX_HIP_TRY(hipGetLastError());
    }
    linompss_exec_opts eo = {1, 128, 128, 0, 30.0, 0};
    int info = 0;
    int rc = linompss_chol_dev(1, 0, n, b, Af, n, s, &eo);
    linompss_exec_stats es;
    linompss_last_exec_stats(&es);
    st->t_factor_s = es.exec_ms * 1e-3;
    if (rc < 0) return rc;
    info = rc;
    bool fallback = info != 0;
    double prev = INFINITY, cur_berr = INFINITY;
    double h_small[1 + 48];
    if (!fallback) {
        // x0 = solve(b) with per-column scaling by ||b_c||_inf
        hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, dB, ldb, n, bmax);
        hipLaunchKernelGGL(lx_d2s_scaled_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, dB, ldb, bmax,
                           Sf);
        rc = linompss_potrs_dev(1, n, nrhs, b, Af, n, Sf, n, s, &eo);
        if (rc) return rc;
        hipLaunchKernelGGL(lx_update_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, Sf, bmax, dX, ldx,
                           0);
        for (int it = 0;; ++it) {
            rc = residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, Y1, Y2, it == 0 ? rowabs : nullptr, colabs, s);
            if (rc) return rc;
            if (it == 0) hipLaunchKernelGGL(lx_max_kernel, dim3(1), dim3(256), 0, s, n, rowabs, anorm);
            hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, R, (int64_t)n, n, rmax);
            hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, dX, ldx, n, xmax);
            hipLaunchKernelGGL(lx_berr_kernel, dim3(1), dim3(64), 0, s, nrhs, rmax, xmax, anorm, berr);
            LX_HIP_TRY(hipGetLastError());
            LX_HIP_TRY(hipMemcpyAsync(h_small, small, sizeof(double) * (1 + 48), hipMemcpyDeviceToHost, s));
            LX_HIP_TRY(hipStreamSynchronize(s));
            double worst = 0.0, rn = 0.0;
            for (int c = 0; c < nrhs; ++c) {
                double e = h_small[1 + 32 + c], r = h_small[1 + c];
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
                               (int64_t)n, rmax, Sf);
            rc = linompss_potrs_dev(1, n, nrhs, b, Af, n, Sf, n, s, &eo);
            if (rc) return rc;
            hipLaunchKernelGGL(lx_update_kernel, dim3(grid_for((int64_t)nr)), dim3(256), 0, s, n, nrhs, Sf, rmax, dX,
                               ldx, 1);
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
        rc = residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, Y1, Y2, rowabs, colabs, s);
        if (rc) return rc;
        hipLaunchKernelGGL(lx_max_kernel, dim3(1), dim3(256), 0, s, n, rowabs, anorm);
        hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, R, (int64_t)n, n, rmax);
        hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, dX, ldx, n, xmax);
        hipLaunchKernelGGL(lx_berr_kernel, dim3(1), dim3(64), 0, s, nrhs, rmax, xmax, anorm, berr);
        LX_HIP_TRY(hipMemcpyAsync(h_small, small, sizeof(double) * (1 + 48), hipMemcpyDeviceToHost, s));
        LX_HIP_TRY(hipStreamSynchronize(s));
        cur_berr = 0.0;
        for (int c = 0; c < nrhs; ++c)
            if (!(h_small[1 + 32 + c] <= cur_berr)) cur_berr = h_small[1 + 32 + c];
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
    if (n <= 0 || nrhs <= 0 || nrhs > 16) return -1;
    std::lock_guard<std::mutex> g(g_ir_mu);
    const size_t nr = (size_t)n * nrhs;
    char *ws = (char *)lx_scratch(7, sizeof(double) * (3 * nr + 2 * (size_t)n + 128));
    if (!ws) return LINOMPSS_ERR_NOMEM;
    double *R = (double *)ws, *Y1 = R + nr, *Y2 = Y1 + nr, *rowabs = Y2 + nr, *colabs = rowabs + n;
    double *small = colabs + n;
    double *anorm = small, *rmax = small + 1, *xmax = small + 1 + 16, *berr = small + 1 + 32;
    hipStream_t s = nullptr;
    int rc = residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, Y1, Y2, rowabs, colabs, s);
    if (rc) return rc;
    hipLaunchKernelGGL(lx_max_kernel, dim3(1), dim3(256), 0, s, n, rowabs, anorm);
    hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, R, (int64_t)n, n, rmax);
    hipLaunchKernelGGL(lx_colmax_kernel, dim3(nrhs), dim3(256), 0, s, dX, ldx, n, xmax);
    hipLaunchKernelGGL(lx_berr_kernel, dim3(1), dim3(64), 0, s, nrhs, rmax, xmax, anorm, berr);
    double h[16];
    LX_HIP_TRY(hipMemcpy(h, berr, sizeof(double) * nrhs, hipMemcpyDeviceToHost));
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c)
        if (!(h[c] <= worst)) worst = h[c];
    *berr_out = worst;
    return 0;
}
