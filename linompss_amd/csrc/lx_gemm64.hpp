// lx_gemm64.hpp -- C = alpha op(A) op(B) + beta C on the matrix cores in
// 64 x 64 output tiles, FP64 or FP32 (v_mfma_{f64,f32}_16x16x4), with an
// optional triangle mask and a deterministic split-K (partials per K chunk,
// then a fixed-order reduction) when the output has too few tiles to fill the
// device.  Shared by the Cholesky adjoint (lx_adjoint.hip) and the standalone
// BLAS-3 entry points (lx_blas3.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lx_device.hpp"

#define LX_G64_T 64   // output tile (rows and columns)
#define LX_G64_KS 16  // k per LDS slab
#define LX_G64_LD (LX_G64_T + 4)

// op(A)[M x K] element (i, k): TA = 0 -> A[i + k lda], TA = 1 -> A[k + i lda].
// op(B)[K x N] element (k, j): TB = 0 -> B[k + j ldb], TB = 1 -> B[j + k ldb].
// ksplit == 1:  C = alpha op(A) op(B) + beta C   (beta == 0: C is not read)
// ksplit  > 1:  part[(z N + j) M + i] = alpha * (partial over the z-th K chunk)
// uplo 1: entries with i < j, uplo 2: entries with i > j are neither read
// nor written (uplo 0: every entry).
template <typename T, int TA, int TB>
__global__ void __launch_bounds__(LX_WG_THREADS)
lx_gemm64_kernel(int M, int N, int K, T alpha, const T *__restrict__ A, int64_t lda, const T *__restrict__ B,
                 int64_t ldb, T beta, T *__restrict__ C, int64_t ldc, int uplo, int ksplit, int kchunk,
                 T *__restrict__ part)
{
    __shared__ T sA[LX_G64_KS][LX_G64_LD], sB[LX_G64_KS][LX_G64_LD];
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    const int m0 = blockIdx.x * LX_G64_T, n0 = blockIdx.y * LX_G64_T, z = blockIdx.z;
    const int kb = z * kchunk, ke = K < kb + kchunk ? K : kb + kchunk;
    // a tile wholly outside the triangle has nothing to do
    if (uplo == 1 && m0 + LX_G64_T - 1 < n0) return;
    if (uplo == 2 && m0 > n0 + LX_G64_T - 1) return;
    // staging: 4 elements per operand per thread per slab
    T ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (TA == 0) {  // rows contiguous
                const int i = m0 + (tid & 63), k = k0 + (tid >> 6) + 4 * e;
                ra[e] = (i < M && k < ke) ? A[(int64_t)k * lda + i] : T(0);
            } else {  // k contiguous
                const int k = k0 + (tid & 15), i = m0 + (tid >> 4) + 16 * e;
                ra[e] = (i < M && k < ke) ? A[(int64_t)i * lda + k] : T(0);
            }
            if (TB == 1) {  // columns contiguous
                const int j = n0 + (tid & 63), k = k0 + (tid >> 6) + 4 * e;
                rb[e] = (j < N && k < ke) ? B[(int64_t)k * ldb + j] : T(0);
            } else {  // k contiguous
                const int k = k0 + (tid & 15), j = n0 + (tid >> 4) + 16 * e;
                rb[e] = (j < N && k < ke) ? B[(int64_t)j * ldb + k] : T(0);
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (TA == 0) sA[(tid >> 6) + 4 * e][tid & 63] = ra[e];
            else sA[tid & 15][(tid >> 4) + 16 * e] = ra[e];
            if (TB == 1) sB[(tid >> 6) + 4 * e][tid & 63] = rb[e];
            else sB[tid & 15][(tid >> 4) + 16 * e] = rb[e];
        }
    };
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    typename lx_mfma<T>::acc acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = typename lx_mfma<T>::acc{T(0), T(0), T(0), T(0)};
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += LX_G64_KS) {
        __syncthreads();  // the previous slab's reads are done
        store();
        __syncthreads();
        if (k0 + LX_G64_KS < ke) load(k0 + LX_G64_KS);  // next slab in flight during the products
#pragma unroll
        for (int kk = 0; kk < LX_G64_KS; kk += 4) {
            T a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = sA[kk + lk][wr + 16 * i + lr];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = sB[kk + lk][wc + 16 * j + lr];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = lx_mfma<T>::mac(a[i], b[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = m0 + wr + 16 * i + lx_mfma<T>::row(lk, e), c = n0 + wc + 16 * j + lr;
                if (r >= M || c >= N || (uplo == 1 && r < c) || (uplo == 2 && r > c)) continue;
                const T v = alpha * acc[i][j][e];
                if (ksplit > 1) {
                    part[((int64_t)z * N + c) * M + r] = v;
                } else {
                    T *p = C + (int64_t)c * ldc + r;
                    *p = beta == T(0) ? v : v + beta * *p;
                }
            }
}

// C = beta C + sum_z part[z], within the uplo triangle.  The partials go into
// eight running sums (z mod 8), so eight loads are in flight per thread, and
// the sums meet in a fixed order: the same result on every run.
template <typename T>
__global__ void lx_gemm64_reduce_kernel(int M, int N, int ksplit, const T *__restrict__ part, T beta,
                                        T *__restrict__ C, int64_t ldc, int uplo)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)M * N) return;
    const int r = (int)(e % M), c = (int)(e / M);
    if ((uplo == 1 && r < c) || (uplo == 2 && r > c)) return;
    const T *p0 = part + (int64_t)c * M + r;
    const int64_t zs = (int64_t)N * M;
    T a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = T(0);
    int z = 0;
    for (; z + 8 <= ksplit; z += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += p0[(z + u) * zs];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (z + u < ksplit) a[u] += p0[(z + u) * zs];
    const T s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    T *p = C + (int64_t)c * ldc + r;
    *p = beta == T(0) ? s : s + beta * *p;
}

#ifndef LX_GEMM64_KERNELS_ONLY
#include <algorithm>

// Launches C = alpha op(A) op(B) + beta C (ta / tb: 0 = N, 1 = T) with split-K
// through part (part_elems values) when M x N has fewer than two tiles per CU
// and K >= 2 min_chunk, in chunks of at least min_chunk (a multiple of 16).
// Returns 0 or LINOMPSS_ERR_HIP.
template <typename T>
int lx_gemm64(int ta, int tb, int M, int N, int K, T alpha, const T *A, int64_t lda, const T *B, int64_t ldb,
              T beta, T *C, int64_t ldc, int uplo, T *part, size_t part_elems, int cus, hipStream_t s,
              int min_chunk = 256)
{
    if (M <= 0 || N <= 0) return 0;
    const int tm = (M + LX_G64_T - 1) / LX_G64_T, tn = (N + LX_G64_T - 1) / LX_G64_T;
    const int target = 2 * cus;
    int ksplit = 1;
    if ((int64_t)tm * tn < target && K >= 2 * min_chunk && part) {
        ksplit = (int)std::min<int64_t>((target + (int64_t)tm * tn - 1) / ((int64_t)tm * tn), K / min_chunk);
        ksplit = (int)std::min<int64_t>(ksplit, (int64_t)(part_elems / ((size_t)M * N)));
        ksplit = std::max(1, ksplit);
    }
    int kchunk = K > 0 ? (K + ksplit - 1) / ksplit : 1;
    kchunk = (kchunk + LX_G64_KS - 1) / LX_G64_KS * LX_G64_KS;
    ksplit = K > 0 ? (K + kchunk - 1) / kchunk : 1;
    const dim3 grid(tm, tn, ksplit);
#define LX_G64_LAUNCH(TA_, TB_)                                                                                      \
    hipLaunchKernelGGL((lx_gemm64_kernel<T, TA_, TB_>), grid, dim3(LX_WG_THREADS), 0, s, M, N, K, alpha, A, lda, B, \
                       ldb, beta, C, ldc, uplo, ksplit, kchunk, part)
    if (ta == 0 && tb == 0) LX_G64_LAUNCH(0, 0);
    else if (ta == 0) LX_G64_LAUNCH(0, 1);
    else if (tb == 0) LX_G64_LAUNCH(1, 0);
    else LX_G64_LAUNCH(1, 1);
#undef LX_G64_LAUNCH
    if (ksplit > 1) {
        const int64_t tot = (int64_t)M * N;
        hipLaunchKernelGGL(lx_gemm64_reduce_kernel<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, M, N,
                           ksplit, part, beta, C, ldc, uplo);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1002;
}
#endif
