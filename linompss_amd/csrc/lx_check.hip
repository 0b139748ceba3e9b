// lx_check.hip -- the factor check of the reference's dchol CLI on the device.
//
// chol_check.c:12-42 forms L2 = L L^T from the lower triangle of the factor
// (lacpy + trmm R/L/T/N) and reports DMAT_RELERR(lower, L2, Aorig)
// (densutil.c:50-99): max_j sum_{i>=j} | |L2_ij| - |A_ij| | divided by
// max_j sum_{i>=j} |L2_ij|.  Here L2 is the lower half of a SYRK on the
// matrix cores over a copy of tril(L), and the column sums are reduced in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "linompss.h"
#include "lx_internal.hpp"

namespace {

template <typename T>
__global__ void lx_tril_copy_kernel(int n, const T *__restrict__ L, int64_t ldl, T *__restrict__ W)
{
    const int64_t tot = (int64_t)n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / n, i = e % n;
        W[e] = i >= j ? L[j * ldl + i] : (T)0;
    }
}

// One workgroup per column: sums of |C_ij| and of | |C_ij| - |A_ij| | over i >= j.
template <typename T>
__global__ void __launch_bounds__(256) lx_relerr_cols_kernel(int n, const T *__restrict__ C, const T *__restrict__ A,
                                                             int64_t lda, double *__restrict__ sums)
{
    __shared__ double sa[256], sd[256];
    const int j = blockIdx.x;
    double a = 0.0, d = 0.0;
    for (int i = j + threadIdx.x; i < n; i += blockDim.x) {
        const double c = fabs((double)C[(int64_t)j * n + i]);
        a += c;
        d += fabs(c - fabs((double)A[(int64_t)j * lda + i]));
    }
    sa[threadIdx.x] = a;
    sd[threadIdx.x] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[2 * j] = sa[0];
        sums[2 * j + 1] = sd[0];
    }
}

template <typename T>
int relerr_impl(int n, const T *dL, int64_t ldl, const T *dA, int64_t lda, double *out)
{
    const size_t nn = (size_t)n * n;
    T *W = nullptr, *C = nullptr;
    double *sums = nullptr;
    int rc = 0;
    if (hipMalloc(&W, sizeof(T) * nn) != hipSuccess || hipMalloc(&C, sizeof(T) * nn) != hipSuccess ||
        hipMalloc(&sums, sizeof(double) * 2 * n) != hipSuccess) {
        rc = LINOMPSS_ERR_NOMEM;
    } else {
        const int grid = (int)std::min<int64_t>(16384, ((int64_t)nn + 255) / 256);
        hipLaunchKernelGGL(lx_tril_copy_kernel<T>, dim3(grid), dim3(256), 0, 0, n, dL, ldl, W);
        rc = hipGetLastError() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
        if (!rc) rc = linompss_syrk_dev(sizeof(T) == 8 ? 0 : 1, 'L', 'N', n, n, 1.0, W, n, 0.0, C, n, nullptr);
        if (!rc) {
            hipLaunchKernelGGL(lx_relerr_cols_kernel<T>, dim3(n), dim3(256), 0, 0, n, C, dA, lda, sums);
            std::vector<double> h(2 * (size_t)n);
            if (hipMemcpy(h.data(), sums, sizeof(double) * 2 * n, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = LINOMPSS_ERR_HIP;
            } else {
                double na = 0.0, nd = 0.0;
                for (int j = 0; j < n; ++j) {
                    na = std::max(na, h[2 * j]);
                    nd = std::max(nd, h[2 * j + 1]);
                }
                *out = nd / na;
            }
        }
    }
    (void)hipFree(W);
    (void)hipFree(C);
    (void)hipFree(sums);
    return rc;
}

}  // namespace

extern "C" int linompss_relerr_lower_dev(int dtype, int n, const void *dL, int64_t ldl, const void *dA,
                                         int64_t lda, double *relerr)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (n < 0) return -2;
    if (ldl < std::max(1, n)) return -4;
    if (lda < std::max(1, n)) return -6;
    if (!relerr) return -7;
    if (n == 0) {
        *relerr = 0.0;
        return 0;
    }
    if (dtype == 0) return relerr_impl<double>(n, (const double *)dL, ldl, (const double *)dA, lda, relerr);
    return relerr_impl<float>(n, (const float *)dL, ldl, (const float *)dA, lda, relerr);
}
