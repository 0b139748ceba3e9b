// lx_gen.hip -- device-side workload generators.
//
// The reference's generators draw from drand48 on the host (GENMAT_SPD,
// src/matutil/genmat.c:49-70).  The benchmark workloads are SPD matrices
// A = G G^T / n + alpha I with G of i.i.d. N(0,1) entries, far too large to
// build on the host, so they are generated here:
//   * G from a counter-based generator, Philox4x32-10 (Salmon et al., SC'11),
//     keyed by the seed and counted by element, then Box-Muller;
//   * A = G G^T / n on the FP64 matrix cores (lower tiles, mirrored);
//   * lambda_max(G G^T / n) by Lanczos, alpha = lambda_max / (kappa - 1)
//     so that cond2(A) = (lambda_max + alpha) / (lambda_min + alpha) ~ kappa
//     (lambda_min of a square Wishart matrix is ~1/n^2, negligible);
//   * right-hand sides uniform in [-1, 1) from the same generator (another key).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "linompss.h"
#include "lx_device.hpp"
#include "lx_internal.hpp"

// Philox4x32-10: 10 rounds of (mulhilo by M0/M1, xor with the key), key
// bumped by the Weyl constants between rounds.
struct lx_u4 {
    uint32_t x, y, z, w;
};
__host__ __device__ inline lx_u4 lx_philox4x32_10(lx_u4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = lx_u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Two 53-bit uniforms in (0, 1] from one Philox block, then Box-Muller.
__host__ __device__ inline void lx_gauss2(uint64_t ctr, uint32_t seed, uint32_t stream, double &g0, double &g1)
{
    lx_u4 r = lx_philox4x32_10(lx_u4{(uint32_t)ctr, (uint32_t)(ctr >> 32), stream, 0u}, seed, 0x1234567u);
    const double u0 = ((double)((((uint64_t)r.x) << 21) ^ (r.y >> 11)) + 1.0) * 0x1p-53;
    const double u1 = ((double)((((uint64_t)r.z) << 21) ^ (r.w >> 11)) + 1.0) * 0x1p-53;
    const double rad = sqrt(-2.0 * log(u0));
    const double th = 6.283185307179586 * u1;
    g0 = rad * cos(th);
    g1 = rad * sin(th);
}

// G[i + j*ldg] for element e = j*n + i (column-major order of generation).
__global__ void lx_gen_gauss_kernel(int m, int n, double *G, int64_t ldg, uint32_t seed)
{
    const int64_t tot = (int64_t)m * n;
    for (int64_t e2 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2; e2 < tot;
         e2 += (int64_t)gridDim.x * blockDim.x * 2) {
        double g0, g1;
        lx_gauss2((uint64_t)(e2 >> 1), seed, 0u, g0, g1);
        G[(e2 / m) * ldg + e2 % m] = g0;
        if (e2 + 1 < tot) G[((e2 + 1) / m) * ldg + (e2 + 1) % m] = g1;
    }
}

template <typename T>
__global__ void lx_gen_uniform_kernel(int m, int n, T *X, int64_t ldx, uint32_t seed, double lo, double hi)
{
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        lx_u4 r = lx_philox4x32_10(lx_u4{(uint32_t)e, (uint32_t)(e >> 32), 1u, 0u}, seed, 0x89abcdefu);
        const double u = (double)((((uint64_t)r.x) << 21) ^ (r.y >> 11)) * 0x1p-53;
        X[(e / m) * ldx + e % m] = (T)fma(hi - lo, u, lo);  // one rounding, then the element type
    }
}

// A = G G^T * scale (+ alpha on the diagonal): one 128 x 128 lower tile per
// workgroup on the FP64 matrix cores, written to both triangles.
__global__ void __launch_bounds__(256) lx_gram_kernel(int n, const double *__restrict__ G, int64_t ldg, int k,
                                                      double *__restrict__ A, int64_t lda, double scale,
                                                      double alpha)
{
    // blockIdx.x enumerates lower tiles (ti >= tj)
    const int t = blockIdx.x;
    int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) / 2.0);
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while (ti * (ti + 1) / 2 > t) --ti;
    const int tj = t - ti * (ti + 1) / 2;
    const int R0 = ti * 128, C0 = tj * 128;
    const int M = std::min(128, n - R0), N = std::min(128, n - C0);
    const int w = threadIdx.x >> 6;
    const int r0 = (w >> 1) * 64, c0 = (w & 1) * 64;
    lx_d4 acc[4][4];
    lx_acc_zero<double, 4, 4>(acc);
    if (r0 < M && c0 < N) lx_wave_gemm_nt<double, 4, 4, 2>(acc, G + R0, ldg, G + C0, ldg, r0, c0, M, N, k);
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + i * 16 + lx_mfma<double>::row(lk, e), c = c0 + j * 16 + lr;
                if (r < M && c < N) {
                    const int gr = R0 + r, gc = C0 + c;
                    double v = acc[i][j][e] * scale + (gr == gc ? alpha : 0.0);
                    A[(int64_t)gc * lda + gr] = v;
                    A[(int64_t)gr * lda + gc] = v;
                }
            }
}

__global__ void lx_add_diag_kernel(int n, double *A, int64_t lda, double alpha)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        A[(int64_t)i * lda + i] += alpha;
}

// y = A x for symmetric A stored full (Lanczos on the generated matrix)
__global__ void __launch_bounds__(256) lx_gemv_full_kernel(int n, const double *__restrict__ A, int64_t lda,
                                                           const double *__restrict__ x, double *__restrict__ y)
{
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = blockIdx.x * 64 + lane;
    const int rr = row < n ? row : n - 1;
    double s = 0.0;
    for (int j = w; j < n; j += 4) s += A[(int64_t)j * lda + rr] * x[j];
    red[w][lane] = s;
    __syncthreads();
    const double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (w == 0 && row < n) y[row] = v;
}

__global__ void __launch_bounds__(256) lx_norm2_kernel(int n, const double *__restrict__ v, double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += v[i] * v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(red[0]);
}

// out[0] = x . y (one workgroup, fixed-order tree: deterministic)
__global__ void __launch_bounds__(256) lx_dot_kernel(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                     double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i] * y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Lanczos three-term step: w -= a v + b u
__global__ void lx_lanczos_axpy_kernel(int n, double *__restrict__ w, const double *__restrict__ v,
                                       const double *__restrict__ u, double a, double b)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        w[i] -= a * v[i] + b * u[i];
}

__global__ void lx_scale_vec_kernel(int n, double *v, const double *nrm)
{
    const double s = nrm[0] > 0 ? 1.0 / nrm[0] : 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] *= s;
}

static int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

extern "C" int linompss_gen_gauss_spd(int n, uint32_t seed, double kappa, double *dA, int64_t lda, double *alpha_out,
                                      double *lmax_out)
{
    if (n <= 0) return -1;
    if (!(kappa > 1.0)) return -3;
    if (lda < n) return -5;
    double *G = nullptr;
    // G, then four Lanczos vectors in its storage
    if (hipMalloc(&G, sizeof(double) * std::max((size_t)n * n, 4 * (size_t)n)) != hipSuccess)
        return LINOMPSS_ERR_NOMEM;
    hipLaunchKernelGGL(lx_gen_gauss_kernel, dim3(grid_for((int64_t)n * n / 2)), dim3(256), 0, 0, n, n, G, (int64_t)n,
                       seed);
    const int nt = (n + 127) / 128;
    const int ntiles = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(lx_gram_kernel, dim3(ntiles), dim3(256), 0, 0, n, G, (int64_t)n, n, dA, lda, 1.0 / n, 0.0);
    if (hipGetLastError() != hipSuccess) {
        (void)hipFree(G);
        return LINOMPSS_ERR_HIP;
    }
    // lambda_max by Lanczos from the vector of ones: the largest Ritz value
    // (from below) converges with the square root of the top spectral gap,
    // which for a Wishart matrix is too small for the power method (60
    // power steps left it 0.75 % low at n = 600).  The Ritz value is the
    // largest eigenvalue of the tridiagonal T_k, found by bisection on its
    // Sturm sequence; iterate until it moves by less than 1e-13 relative.
    double *v = G, *u = G + n, *w = G + 2 * (size_t)n, *sc = G + 3 * (size_t)n;  // reuse G's storage
    (void)hipDeviceSynchronize();
    {
        std::vector<double> ones(n, 1.0 / std::sqrt((double)n));
        if (hipMemcpy(v, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(u, 0, sizeof(double) * n) != hipSuccess) {
            (void)hipFree(G);
            return LINOMPSS_ERR_HIP;
        }
    }
    std::vector<double> al, be;  // T_k: diagonal al[0..k-1], off-diagonal be[1..k-1]
    be.push_back(0.0);
    auto ritz_max = [&]() {
        // number of eigenvalues of T below x (Sturm count), then bisection
        const int k = (int)al.size();
        double lo = al[0], hi = al[0];
        for (int i = 0; i < k; ++i) {
            const double r = (i > 0 ? std::fabs(be[i]) : 0.0) + (i + 1 < k ? std::fabs(be[i + 1]) : 0.0);
            lo = std::min(lo, al[i] - r);
            hi = std::max(hi, al[i] + r);
        }
        auto below = [&](double x) {
            int c = 0;
            double d = 1.0;
            for (int i = 0; i < k; ++i) {
                d = al[i] - x - (i > 0 ? be[i] * be[i] / d : 0.0);
                if (d == 0.0) d = -1e-300;
                if (d < 0.0) ++c;
            }
            return c;
        };
        for (int it = 0; it < 200 && hi - lo > 1e-15 * std::fabs(hi); ++it) {
            const double mid = 0.5 * (lo + hi);
            if (below(mid) >= k) hi = mid;  // every eigenvalue below mid
            else lo = mid;
        }
        return hi;
    };
    double lam = 0.0;
    for (int j = 0; j < 400; ++j) {
        hipLaunchKernelGGL(lx_gemv_full_kernel, dim3((n + 63) / 64), dim3(256), 0, 0, n, dA, lda, v, w);
        hipLaunchKernelGGL(lx_dot_kernel, dim3(1), dim3(256), 0, 0, n, w, v, sc);
        double a = 0.0;
        if (hipMemcpy(&a, sc, sizeof a, hipMemcpyDeviceToHost) != hipSuccess) {
            (void)hipFree(G);
            return LINOMPSS_ERR_HIP;
        }
        hipLaunchKernelGGL(lx_lanczos_axpy_kernel, dim3(grid_for(n)), dim3(256), 0, 0, n, w, v, u, a, be.back());
        hipLaunchKernelGGL(lx_norm2_kernel, dim3(1), dim3(256), 0, 0, n, w, sc);
        double bn = 0.0;
        if (hipMemcpy(&bn, sc, sizeof bn, hipMemcpyDeviceToHost) != hipSuccess) {
            (void)hipFree(G);
            return LINOMPSS_ERR_HIP;
        }
        al.push_back(a);
        const double prev = lam;
        if (j % 5 == 4 || !(bn > 0.0)) {
            lam = ritz_max();
            if (j >= 9 && std::fabs(lam - prev) <= 1e-13 * lam) break;
        }
        if (!(bn > 1e-300 * std::fabs(a))) break;  // an invariant subspace: T_k is exact
        be.push_back(bn);
        hipLaunchKernelGGL(lx_scale_vec_kernel, dim3(grid_for(n)), dim3(256), 0, 0, n, w, sc);
        std::swap(u, v);  // u = v_j
        std::swap(v, w);  // v = v_{j+1}
    }
    if (al.size() % 5 != 0) lam = std::max(lam, ritz_max());
    const double alpha = lam / (kappa - 1.0);
    hipLaunchKernelGGL(lx_add_diag_kernel, dim3(grid_for(n)), dim3(256), 0, 0, n, dA, lda, alpha);
    const hipError_t e = hipDeviceSynchronize();
    (void)hipFree(G);
    if (e != hipSuccess) return LINOMPSS_ERR_HIP;
    if (alpha_out) *alpha_out = alpha;
    if (lmax_out) *lmax_out = lam;
    return 0;
}

extern "C" int linompss_gen_uniform(int dtype, int m, int n, uint32_t seed, double lo, double hi, void *dX,
                                    int64_t ldx)
{
    if (m < 0 || n < 0) return -2;
    if (ldx < std::max(1, m)) return -8;
    if ((int64_t)m * n == 0) return 0;
    if (dtype == 0)
        hipLaunchKernelGGL(lx_gen_uniform_kernel<double>, dim3(grid_for((int64_t)m * n)), dim3(256), 0, 0, m, n,
                           (double *)dX, ldx, seed, lo, hi);
    else
        hipLaunchKernelGGL(lx_gen_uniform_kernel<float>, dim3(grid_for((int64_t)m * n)), dim3(256), 0, 0, m, n,
                           (float *)dX, ldx, seed, lo, hi);
    return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

extern "C" int linompss_gen_gauss(int m, int n, uint32_t seed, double *dG, int64_t ldg)
{
    if (m < 0 || n < 0) return -1;
    if ((int64_t)m * n == 0) return 0;
    hipLaunchKernelGGL(lx_gen_gauss_kernel, dim3(grid_for((int64_t)m * n / 2 + 1)), dim3(256), 0, 0, m, n, dG, ldg,
                       seed);
    return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

// The refinement app's default matrix, GENMAT_COVARIANT (src/matutil/
// genmat.c:263-302, called from itref_setup.c:115 through dgenmat_config with
// p = k = 2, genmat_config.c:37): a_ij = |i - j|^-p off the diagonal,
// a_jj = 1 + (j + 1)^(1/k); both triangles written.  One thread per element,
// rows of a column on consecutive lanes.
__global__ void lx_gen_covariant_kernel(int n, double mp, double kinv, double *__restrict__ A, int64_t lda)
{
    const int64_t tot = (int64_t)n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / n), i = (int)(e % n);
        A[(int64_t)j * lda + i] = i == j ? 1.0 + pow((double)(j + 1), kinv) : pow((double)(i > j ? i - j : j - i), mp);
    }
}

extern "C" int linompss_gen_covariant(int n, double p, double k, double *dA, int64_t lda)
{
    if (n < 0) return -1;
    if (!(p > 0.0)) return -2;
    if (!(k != 0.0)) return -3;
    if (lda < std::max(1, n)) return -5;
    if (n == 0) return 0;
    hipLaunchKernelGGL(lx_gen_covariant_kernel, dim3(grid_for((int64_t)n * n)), dim3(256), 0, 0, n, -p, 1.0 / k, dA,
                       lda);
    return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

// ---- The reference's own generators, GENMAT_SPD and GENMAT
// (src/matutil/genmat.c:49-70, 106-128), for its CLI front ends: drand48
// (x <- 0x5DEECE66D x + 0xB mod 2^48, value x / 2^48, seeded x0 = seed << 16
// | 0x330E) drawn in the reference's loop order.  Each element jumps to its
// own draw: with x_s = A_s x0 + C_s, the jumps for s = 2^b are tabulated on
// the host (A_2s = A_s^2, C_2s = (A_s + 1) C_s) and composed over the set bits
// of s, all modulo 2^64 and masked to 48 bits at the end (2^48 divides 2^64).
struct lx_lcg_table {
    uint64_t a[48], c[48];
};

static lx_lcg_table lx_lcg_make()
{
    lx_lcg_table t;
    uint64_t a = 0x5DEECE66Dull, c = 0xBull;
    for (int b = 0; b < 48; ++b) {
        t.a[b] = a;
        t.c[b] = c;
        c = (a + 1) * c;
        a = a * a;
    }
    return t;
}

__device__ inline double lx_drand48_at(const lx_lcg_table &t, uint64_t x0, uint64_t s)
{
    uint64_t x = x0;
    for (int b = 0; s; ++b, s >>= 1)
        if (s & 1) x = t.a[b] * x + t.c[b];
    return ldexp((double)(x & 0xFFFFFFFFFFFFull), -48);
}

// GENMAT_SPD: column j draws rows i = j..n-1 in order, the diagonal gets + n
// (in fp_t arithmetic), the value is mirrored to (j, i).  Element (i, j) of
// the lower triangle is draw number j n - j (j - 1) / 2 + (i - j) + 1.
template <typename T>
__global__ void lx_genmat_spd_kernel(lx_lcg_table t, uint64_t x0, int n, T *__restrict__ A, int64_t lda)
{
    const int64_t tot = (int64_t)n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / n, i = e % n;
        const int64_t c = i >= j ? j : i, r = i >= j ? i : j;
        const uint64_t s = (uint64_t)(c * n - c * (c - 1) / 2 + (r - c) + 1);
        T v = (T)lx_drand48_at(t, x0, s);
        if (i == j) v += (T)n;
        A[j * lda + i] = v;
    }
}

// GENMAT: X(i, j) is draw number j m + i + 1, converted to fp_t.
template <typename T>
__global__ void lx_genmat_kernel(lx_lcg_table t, uint64_t x0, int m, int n, T *__restrict__ X, int64_t ldx)
{
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x)
        X[(e / m) * ldx + e % m] = (T)lx_drand48_at(t, x0, (uint64_t)e + 1);
}

static uint64_t lx_drand48_seed(long seed) { return ((uint64_t)(uint32_t)seed << 16) | 0x330Eu; }

extern "C" int linompss_gen_drand48_spd(int dtype, int n, long seed, void *dA, int64_t lda)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (n < 0) return -2;
    if (lda < std::max(1, n)) return -5;
    if (n == 0) return 0;
    const lx_lcg_table t = lx_lcg_make();
    const dim3 g(grid_for((int64_t)n * n)), b(256);
    if (dtype == 0)
        hipLaunchKernelGGL(lx_genmat_spd_kernel<double>, g, b, 0, 0, t, lx_drand48_seed(seed), n, (double *)dA, lda);
    else
        hipLaunchKernelGGL(lx_genmat_spd_kernel<float>, g, b, 0, 0, t, lx_drand48_seed(seed), n, (float *)dA, lda);
    return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

extern "C" int linompss_gen_drand48(int dtype, int m, int n, long seed, void *dX, int64_t ldx)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (ldx < std::max(1, m)) return -6;
    if ((int64_t)m * n == 0) return 0;
    const lx_lcg_table t = lx_lcg_make();
    const dim3 g(grid_for((int64_t)m * n)), b(256);
    if (dtype == 0)
        hipLaunchKernelGGL(lx_genmat_kernel<double>, g, b, 0, 0, t, lx_drand48_seed(seed), m, n, (double *)dX, ldx);
    else
        hipLaunchKernelGGL(lx_genmat_kernel<float>, g, b, 0, 0, t, lx_drand48_seed(seed), m, n, (float *)dX, ldx);
    return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}
