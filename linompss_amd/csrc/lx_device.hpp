// lx_device.hpp -- gfx950 tile kernels of the tiled Cholesky path.
//
// Every routine here is a *work unit body*: it is executed by one workgroup of
// LX_WG_THREADS (4 wavefronts of 64 lanes) inside the persistent executor
// (lx_exec.hip).  Matrices are column-major with 64-bit leading dimensions,
// exactly the layout the reference hands to its tile tasks
// (src/blocked_blas/tasks_*.c, one LAPACK/BLAS call per b x b tile).
//
// Arithmetic: the trailing updates (reference task_dgemm / task_dsyrk,
// src/blocked_blas/task_gemm.c:55-58, tasks_syrk.c:19-21) run on the matrix
// cores with v_mfma_f64_16x16x4_f64 (FP64) or v_mfma_f32_16x16x4_f32 (FP32);
// the diagonal factorisation (task_dpotrf, tasks_potrf.c:20-24) and the
// triangular solves (task_dtrsm, tasks_trsm.c:22-25) use wavefront-level code
// with their inner updates on the same MFMA core.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LX_WG_THREADS 256

typedef double lx_d4 __attribute__((ext_vector_type(4)));
typedef float lx_f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// MFMA 16x16x4 for both precisions.  Operand layout (identical for f64 and
// f32): lane l supplies A[row = l%16][k = l/16] and B[k = l/16][col = l%16].
// The 16x16 result is held as 4 values per lane in column l%16; the rows
// differ (measured on gfx950 with an in-bounds probe of each builtin):
//   f32: value i of lane l is row 4*(l/16) + i
//   f64: value i of lane l is row (l/16) + 4*i
template <typename T> struct lx_mfma;
template <> struct lx_mfma<double> {
    typedef lx_d4 acc;
    static __device__ __forceinline__ acc mac(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int i) { return lk + 4 * i; }
};
template <> struct lx_mfma<float> {
    typedef lx_f4 acc;
    static __device__ __forceinline__ acc mac(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int i) { return 4 * lk + i; }
};

// Thread index passed through an empty asm statement.  The executor runs
// many unit bodies inside one persistent loop; with the plain index the
// compiler hoists every lane-derived address of every body out of that loop
// and keeps them all live at once (hundreds of registers, then spills).  The
// opaque copy is recomputed where it is used.
__device__ __forceinline__ int lx_tid()
{
    int t = threadIdx.x;
#if defined(__AMDGCN__)
    asm volatile("" : "+v"(t));
#endif
    return t;
}
__device__ __forceinline__ int lx_lane() { return lx_tid() & 63; }
__device__ __forceinline__ int lx_wave() { return threadIdx.x >> 6; }

// ---------------------------------------------------------------------------
// Wave-level C_acc += A * B^T over k in [0, K).
//   A element (r, k) at A[k*lda + r], rows r in [r0, r0 + 16*FM)
//   B element (c, k) at B[k*ldb + c], cols c in [c0, c0 + 16*FN)
// Rows >= M and cols >= N are clamped to the last valid one: their products
// land in accumulator entries the caller never stores.  Columns k >= K read a
// clamped, in-bounds address and the A operand is zeroed there.
// Addressing: one uniform base pointer per operand that advances 4*U columns
// per step, plus a 32-bit per-lane element offset (a unit's operand spans at
// most 4 columns ahead of the base, so offsets stay far below 2^31).  Loads of
// the next K-slab are issued before the MFMAs of the current one (register
// ping-pong), so global-load latency hides behind the matrix core.
template <typename T, int FM, int FN, int U>
__device__ __forceinline__ void lx_wave_gemm_nt(typename lx_mfma<T>::acc (&acc)[FM][FN],
                                                const T *__restrict__ A, long long lda,
                                                const T *__restrict__ B, long long ldb,
                                                int r0, int c0, int M, int N, int K)
{
    if (K <= 0 || M <= 0 || N <= 0) return;
    const int lane = lx_lane();
    const int lr = lane & 15, lk = lane >> 4;
    uint32_t oa[FM], ob[FN];  // per-lane byte offsets from the slab base
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        int r = r0 + i * 16 + lr;
        oa[i] = (uint32_t)(((r < M ? r : M - 1) + lk * lda) * (long long)sizeof(T));
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        int c = c0 + j * 16 + lr;
        ob[j] = (uint32_t)(((c < N ? c : N - 1) + lk * ldb) * (long long)sizeof(T));
    }
    constexpr int KS = 4 * U;
    const int kfull = K - K % KS;  // columns covered by unmasked steps

    auto load = [&](T(&as)[U][FM], T(&bs)[U][FN], int k0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const char *Ak = (const char *)(A + (long long)(k0 + 4 * u) * lda);
            const char *Bk = (const char *)(B + (long long)(k0 + 4 * u) * ldb);
#pragma unroll
            for (int i = 0; i < FM; ++i) as[u][i] = *(const T *)(Ak + oa[i]);
#pragma unroll
            for (int j = 0; j < FN; ++j) bs[u][j] = *(const T *)(Bk + ob[j]);
        }
    };
    auto compute = [&](const T(&as)[U][FM], const T(&bs)[U][FN]) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = lx_mfma<T>::mac(as[u][i], bs[u][j], acc[i][j]);
    };

    if (kfull > 0) {
        T ca[U][FM], cb[U][FN];
        load(ca, cb, 0);
        for (int k0 = 0; k0 < kfull; k0 += KS) {
            // prefetch the next slab (the last iteration re-reads its own slab: always in bounds)
            const int kn = k0 + KS < kfull ? k0 + KS : k0;
            T na[U][FM], nb[U][FN];
            load(na, nb, kn);
            compute(ca, cb);
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int i = 0; i < FM; ++i) ca[u][i] = na[u][i];
#pragma unroll
                for (int j = 0; j < FN; ++j) cb[u][j] = nb[u][j];
            }
        }
    }
    // ragged tail: steps of 4 columns, out-of-range k masked on the A operand
    for (int k0 = kfull; k0 < K; k0 += 4) {
        const bool kv = k0 + lk < K;
        const int kb = kv ? k0 : K - 1 - lk;  // slab base keeping this lane's column in range
        const char *Ak = (const char *)(A + (long long)kb * lda);
        const char *Bk = (const char *)(B + (long long)kb * ldb);
        T as[FM], bs[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            T v = *(const T *)(Ak + oa[i]);
            as[i] = kv ? v : T(0);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) bs[j] = *(const T *)(Bk + ob[j]);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = lx_mfma<T>::mac(as[i], bs[j], acc[i][j]);
    }
}

template <typename T, int FM, int FN>
__device__ __forceinline__ void lx_acc_zero(typename lx_mfma<T>::acc (&acc)[FM][FN])
{
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[i][j][e] = T(0);
}

// C[r][c] -= acc for r < M, c < N (and r >= c + diag when lower).  r0/c0 are
// the wave block's origin inside C; gdiag is (global row - global col) of
// C's element (0,0), so the lower mask keeps r + gdiag >= c.
template <typename T, int FM, int FN>
__device__ __forceinline__ void lx_acc_sub_store(typename lx_mfma<T>::acc (&acc)[FM][FN], T *__restrict__ C,
                                                 long long ldc, int r0, int c0, int M, int N, bool lower,
                                                 int gdiag)
{
    const int lane = lx_lane();
    const int lr = lane & 15, lk = lane >> 4;
    // one row of tiles at a time, all its loads before any of its stores: a
    // store may alias a later load as far as the compiler knows, so load/store
    // pairs written interleaved would each wait a full memory round trip
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        T v[FN][4];
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int c = c0 + j * 16 + lr, cc = c < N ? c : N - 1;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + i * 16 + lx_mfma<T>::row(lk, e), rr = r < M ? r : M - 1;
                v[j][e] = C[(long long)cc * ldc + rr];
            }
        }
        // keep the row's loads together where the kernel is short of registers
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int c = c0 + j * 16 + lr;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + i * 16 + lx_mfma<T>::row(lk, e);
                if (r < M && c < N && (!lower || r + gdiag >= c)) C[(long long)c * ldc + r] = v[j][e] - acc[i][j][e];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// ---------------------------------------------------------------------------
// GEMM / SYRK trailing-update unit:  C[M x N] -= A[M x K] * B[N x K]^T  (lower
// mask on the diagonal of a SYRK tile).  The WG covers up to
// (2*16*FM) x (2*16*FN) of C with its 2 x 2 wavefronts.  Wave blocks that lie
// strictly above the diagonal of a SYRK tile are skipped.
template <typename T, int FM, int FN, int U>
__device__ __forceinline__ void lx_unit_gemm_nt(T *__restrict__ C, long long ldc, const T *__restrict__ A,
                                                long long lda, const T *__restrict__ B, long long ldb, int M,
                                                int N, int K, bool lower, int gdiag)
{
    const int w = lx_wave();
    const int r0 = (w >> 1) * 16 * FM, c0 = (w & 1) * 16 * FN;
    if (r0 >= M || c0 >= N) return;
    if (lower && r0 + 16 * FM - 1 + gdiag < c0) return;  // block strictly above the diagonal
    typename lx_mfma<T>::acc acc[FM][FN];
    lx_acc_zero<T, FM, FN>(acc);
    lx_wave_gemm_nt<T, FM, FN, U>(acc, A, lda, B, ldb, r0, c0, M, N, K);
    lx_acc_sub_store<T, FM, FN>(acc, C, ldc, r0, c0, M, N, lower, gdiag);
}

// ---------------------------------------------------------------------------
// LDS-staged GEMM / SYRK unit for the trailing updates (same contract as
// lx_unit_gemm_nt, M, N <= 128):  C[M x N] -= A[M x K] * B[N x K]^T.
// The workgroup stages K-slabs of A and B (128 x KS each) in LDS, double
// buffered: the global loads of slab s+1 (16-byte vectors when the unit is
// full) are in flight while the matrix cores consume slab s, and each operand
// element crosses the memory hierarchy once per workgroup instead of once per
// wave.  LDS layout [k][row] with a row stride of 144 elements, so the four
// 16-lane groups of an MFMA operand read hit disjoint banks.
template <typename T> struct lx_lgemm_cfg;
template <> struct lx_lgemm_cfg<float> { static constexpr int KS = 32; };
template <> struct lx_lgemm_cfg<double> { static constexpr int KS = 16; };
#define LX_LG_STR 144
template <typename T>
struct lx_lgemm_smem {
    static constexpr int bytes = 2 * 2 * lx_lgemm_cfg<T>::KS * LX_LG_STR * (int)sizeof(T);
};

typedef float lx_f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double lx_d2u __attribute__((ext_vector_type(2), aligned(8)));
template <typename T> struct lx_vec16;
template <> struct lx_vec16<float> { typedef lx_f4u type; static constexpr int n = 4; };
template <> struct lx_vec16<double> { typedef lx_d2u type; static constexpr int n = 2; };

template <typename T>
__device__ __forceinline__ void lx_unit_gemm_lds(T *__restrict__ C, long long ldc, const T *__restrict__ A,
                                                 long long lda, const T *__restrict__ B, long long ldb, int M, int N,
                                                 int K, bool lower, int gdiag, char *smem)
{
    constexpr int KS = lx_lgemm_cfg<T>::KS, VN = lx_vec16<T>::n;
    typedef typename lx_vec16<T>::type V;
    constexpr int NV = 128 * KS / VN / LX_WG_THREADS;  // 16-byte vectors per thread per operand and slab
    constexpr int VPC = 128 / VN;                      // vectors per slab column
    T(*sA)[KS][LX_LG_STR] = reinterpret_cast<T(*)[KS][LX_LG_STR]>(smem);
    T(*sB)[KS][LX_LG_STR] = reinterpret_cast<T(*)[KS][LX_LG_STR]>(smem + 2 * KS * LX_LG_STR * sizeof(T));
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
    // a wave block strictly above the diagonal of a SYRK tile does no arithmetic
    const bool active = wr < M && wc < N && !(lower && wr + 63 + gdiag < wc);
    const bool full = M == 128 && N == 128 && K % KS == 0;  // uniform
    const int nslab = (K + KS - 1) / KS;
    typename lx_mfma<T>::acc acc[4][4];
    lx_acc_zero<T, 4, 4>(acc);
    auto compute = [&](int buf) {
        if (!active) return;
        // two k-steps of operands at a time (the barrier stops the compiler from
        // reading the whole slab's operands into registers up front)
#pragma unroll
        for (int k2 = 0; k2 < KS; k2 += 8) {
#pragma unroll
            for (int kk = k2; kk < k2 + 8; kk += 4) {
                T a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = sA[buf][kk + lk][wr + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = sB[buf][kk + lk][wc + j * 16 + lr];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = lx_mfma<T>::mac(a[i], b[j], acc[i][j]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    if (full) {
        // 16-byte loads of slab s+1 in flight while slab s is multiplied
        V ra[NV], rb[NV];
        auto gload = [&](int k0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int q = tid + i * LX_WG_THREADS, kk = q / VPC, r = (q % VPC) * VN;
                ra[i] = *(const V *)(A + (long long)(k0 + kk) * lda + r);
                rb[i] = *(const V *)(B + (long long)(k0 + kk) * ldb + r);
            }
        };
        auto lstore = [&](int buf) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int q = tid + i * LX_WG_THREADS, kk = q / VPC, r = (q % VPC) * VN;
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    sA[buf][kk][r + e] = ra[i][e];
                    sB[buf][kk][r + e] = rb[i][e];
                }
            }
        };
        gload(0);
        lstore(0);
        __syncthreads();
        for (int sl = 0; sl < nslab; ++sl) {
            const int buf = sl & 1;
            if (sl + 1 < nslab) gload((sl + 1) * KS);
            compute(buf);
            if (sl + 1 < nslab) lstore(buf ^ 1);
            __syncthreads();
        }
    } else {
        // ragged unit (edge tiles): element loads straight into LDS, rows past
        // M / N clamped (never stored), columns past K zero; a constant trip
        // count, unrolled, so the loads of a slab are in flight together
        for (int sl = 0; sl < nslab; ++sl) {
            const int k0 = sl * KS;
#pragma unroll 8
            for (int q = 0; q < 128 * KS / LX_WG_THREADS; ++q) {
                const int e = tid + q * LX_WG_THREADS, r = e % 128, kk = e / 128;
                const bool kv = k0 + kk < K;
                const int kc = kv ? k0 + kk : K - 1;
                const T va = A[(long long)kc * lda + (r < M ? r : M - 1)];
                const T vb = B[(long long)kc * ldb + (r < N ? r : N - 1)];
                sA[0][kk][r] = kv ? va : T(0);
                sB[0][kk][r] = kv ? vb : T(0);
            }
            __syncthreads();
            compute(0);
            __syncthreads();
        }
    }
    if (active) lx_acc_sub_store<T, 4, 4>(acc, C, ldc, wr, wc, M, N, lower, gdiag);
}

// ---------------------------------------------------------------------------
// Trailing update of the FP32 factor on the BF16 matrix cores, split operands:
// every FP32 input x is written as hi + mid + lo, three BF16 numbers (8
// significant bits each, the FP32 exponent range, so no scaling is needed),
// and A B^T is accumulated in FP32 from the six products whose size is at
// least 2^-16 of the leading one (lo*hi, hi*lo, mid*mid, mid*hi, hi*mid,
// hi*hi, smallest first): the dropped terms (mid*lo, lo*mid, lo*lo) are below
// 2^-24 relative, so the update carries FP32 accuracy at the BF16 matrix rate
// (2.5 PFLOP/s dense, six products = 16x the MFMA work of one FP32 pass at
// 1/16 of its cost per flop).
//
// Same contract as lx_unit_gemm_lds (M, N <= 128, T = float).  K-slabs of 32:
// each thread loads 16 consecutive k of one row of A and of B, splits them in
// registers and writes the three planes to LDS as [row][k] (row stride 40
// elements: the eight rows of a 128-byte LDS access fall on disjoint banks);
// the loads of slab s+1 are in flight while slab s is multiplied.  Waves own
// 64 x 64 of C as 2 x 2 tiles of v_mfma_f32_32x32x16_bf16 (measured gfx950
// layout: lane l supplies A[l%32][8(l/32)+e], B[8(l/32)+e][l%32], e < 8;
// result value i of lane l is row 8(i/4) + 4(l/32) + i%4, column l%32).
typedef __bf16 lx_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lx_f16v __attribute__((ext_vector_type(16)));
#define LX_BX_KS 32
#define LX_BX_STR 40
struct lx_bx_smem {
    static constexpr int bytes = 2 * 3 * 128 * LX_BX_STR * 2;  // two operands, three planes
};

__device__ __forceinline__ void lx_bx_split8(const float (&x)[8], lx_bf16x8 &h, lx_bf16x8 &m, lx_bf16x8 &l)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 bh = (__bf16)x[e];
        const float r1 = x[e] - (float)bh;
        const __bf16 bm = (__bf16)r1;
        const float r2 = r1 - (float)bm;
        h[e] = bh;
        m[e] = bm;
        l[e] = (__bf16)r2;
    }
}

// The two split kinds of the bodies below.
//   H16 = false: three BF16 planes (hi, mid, lo), six products (above).
//   H16 = true:  two FP16 planes of x * 2^12: hi = fp16(x'), lo = fp16(x' - hi)
//   (x' - hi is exact in FP32), so x' = hi + lo within 2^-22 |x'|; three
//   products (lo*hi, hi*lo, hi*hi; the dropped lo*lo is below 2^-22), each
//   exact in FP32, accumulated in FP32 and scaled back by 2^-24.  Half the
//   matrix-core work and two thirds of the LDS planes of the BF16 split, at
//   about 4x its rounding error per product (still below FP32's own
//   accumulation error over a 512-deep K).  FP16's range needs |x| <= 16
//   (x' <= 65504); the refinement driver equilibrates A to a unit diagonal
//   (powers of two) so every entry of L is at most 1.  Past the range the hi
//   part is infinite and the factor reports a non-positive pivot.  The lo part
//   of |x| < 2^-15 is subnormal: its absolute error stays below 2^-37.
typedef _Float16 lx_f16x8 __attribute__((ext_vector_type(8)));
template <bool H16> struct lx_split;
template <> struct lx_split<false> {
    typedef __bf16 e;
    typedef lx_bf16x8 v;
    static constexpr int np = 3, nq = 6;
    static constexpr int pa(int q) { return q == 0 ? 2 : q == 1 ? 0 : q < 4 ? 1 : 0; }  // 2 0 1 1 0 0
    static constexpr int pb(int q) { return q == 0 ? 0 : q == 1 ? 2 : q == 2 ? 1 : q == 3 ? 0 : q == 4 ? 1 : 0; }
    __device__ static __forceinline__ void split(const float (&x)[8], v (&p)[3]) { lx_bx_split8(x, p[0], p[1], p[2]); }
    __device__ static __forceinline__ lx_f16v mma(v a, v b, lx_f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ float unscale(float a) { return a; }
};
template <> struct lx_split<true> {
    typedef _Float16 e;
    typedef lx_f16x8 v;
    static constexpr int np = 2, nq = 3;
    static constexpr int pa(int q) { return q == 0 ? 1 : 0; }  // 1 0 0
    static constexpr int pb(int q) { return q == 1 ? 1 : 0; }  // 0 1 0
    __device__ static __forceinline__ void split(const float (&x)[8], v (&p)[3])
    {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float xs = x[k] * 4096.0f;
            const _Float16 h = (_Float16)xs;
            p[0][k] = h;
            p[1][k] = (_Float16)(xs - (float)h);
        }
    }
    __device__ static __forceinline__ lx_f16v mma(v a, v b, lx_f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ float unscale(float a) { return a * 0x1p-24f; }
};

template <bool H16 = false>
__device__ __forceinline__ void lx_unit_gemm_bx6(float *__restrict__ C, long long ldc, const float *__restrict__ A,
                                                 long long lda, const float *__restrict__ B, long long ldb, int M,
                                                 int N, int K, bool lower, int gdiag, char *smem)
{
    typedef lx_split<H16> S;
    typedef typename S::e E;
    typedef typename S::v V;
    constexpr int KS = LX_BX_KS, STR = LX_BX_STR;
    // planes: [operand][plane][row][k]
    E(*sP)[3][128][STR] = reinterpret_cast<E(*)[3][128][STR]>(smem);
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave();
    const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
    const bool active = wr < M && wc < N && !(lower && wr + 63 + gdiag < wc);
    const int nslab = (K + KS - 1) / KS;
    // staging: thread -> (row, 16 consecutive k), rows clamped (never stored), k past K zero
    const int srow = tid & 127, sk = (tid >> 7) * 16;
    const float *pa = A + (srow < M ? srow : M - 1);
    const float *pb = B + (srow < N ? srow : N - 1);
    // two register sets: the loads of slab s+2 are in flight while slab s is
    // multiplied and slab s+1 waits in the other set
    float xa[2][16], xb[2][16];
    const bool kfull = K % KS == 0;  // uniform
    auto gload = [&](int set, int k0) {
        if (kfull) {
            // plain loads: their results are first used two slabs later
            const float *qa = pa + (long long)(k0 + sk) * lda, *qb = pb + (long long)(k0 + sk) * ldb;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                xa[set][j] = qa[(long long)j * lda];
                xb[set][j] = qb[(long long)j * ldb];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int k = k0 + sk + j;
                const bool kv = k < K;
                const long long kc = kv ? k : K - 1;
                const float va = pa[kc * lda], vb = pb[kc * ldb];
                xa[set][j] = kv ? va : 0.f;
                xb[set][j] = kv ? vb : 0.f;
            }
        }
    };
    auto lstore = [&](int set) {
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            float ta[8], tb[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                ta[e] = xa[set][half * 8 + e];
                tb[e] = xb[set][half * 8 + e];
            }
            V pl[3];
            S::split(ta, pl);
#pragma unroll
            for (int q = 0; q < S::np; ++q) *(V *)&sP[0][q][srow][sk + half * 8] = pl[q];
            S::split(tb, pl);
#pragma unroll
            for (int q = 0; q < S::np; ++q) *(V *)&sP[1][q][srow][sk + half * 8] = pl[q];
        }
    };
    lx_f16v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;
    const int lr = lane & 31, lk = lane >> 5;
    auto compute = [&]() {
        if (!active) return;
#pragma unroll
        for (int ks = 0; ks < KS; ks += 16) {
            V a[2][S::np], b[2][S::np];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int p = 0; p < S::np; ++p) a[i][p] = *(const V *)&sP[0][p][wr + i * 32 + lr][ks + 8 * lk];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int p = 0; p < S::np; ++p) b[j][p] = *(const V *)&sP[1][p][wc + j * 32 + lr][ks + 8 * lk];
            // product-major: the four tiles' accumulators are independent, so
            // consecutive MFMAs never wait on each other's result
#pragma unroll
            for (int q = 0; q < S::nq; ++q)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = S::mma(a[i][S::pa(q)], b[j][S::pb(q)], acc[i][j]);
        }
    };
    gload(0, 0);
    // the previous block of the same unit may still be reading the planes
    __syncthreads();
    lstore(0);
    __syncthreads();
    if (nslab > 1) gload(0, KS);
    if (nslab > 2) gload(1, 2 * KS);
    // the loop is unrolled by two so that the register sets have fixed names:
    // on entry slab sl is in LDS, slab sl+1 in set 0 and slab sl+2 in set 1
    for (int sl = 0; sl < nslab; sl += 2) {
        compute();  // slab sl
        if (sl + 1 < nslab) {
            __syncthreads();
            lstore(0);  // slab sl+1
            __syncthreads();
            if (sl + 3 < nslab) gload(0, (sl + 3) * KS);
            compute();  // slab sl+1
            if (sl + 2 < nslab) {
                __syncthreads();
                lstore(1);  // slab sl+2
                __syncthreads();
                if (sl + 4 < nslab) gload(1, (sl + 4) * KS);
            }
        }
    }
    if (active) {
        // C -= acc, one row of tiles at a time, its loads before its stores
        const int lane2 = lx_lane(), lr2 = lane2 & 31, lk2 = lane2 >> 5;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            float v[2][16];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int c = wc + j * 32 + lr2, cc = c < N ? c : N - 1;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int r = wr + i * 32 + 8 * (e / 4) + 4 * lk2 + (e % 4), rr = r < M ? r : M - 1;
                    v[j][e] = C[(long long)cc * ldc + rr];
                }
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int c = wc + j * 32 + lr2;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int r = wr + i * 32 + 8 * (e / 4) + 4 * lk2 + (e % 4);
                    if (r < M && c < N && (!lower || r + gdiag >= c))
                        C[(long long)c * ldc + r] = v[j][e] - S::unscale(acc[i][j][e]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// The same split-BF16 update on a 128 x 256 block (M <= 128, N <= 256): waves
// own 64 x 128 of C (2 x 4 tiles of 32 x 32), so each staged element of A
// feeds twice the MFMAs of the 128 x 128 unit and the staging (global loads,
// split, LDS stores) per product drops to three quarters.  K-slabs of 16 to
// keep the planes within 54 KB (two workgroups per CU): planes [plane][row][k]
// with rows 0..127 = A and 128..383 = B, row stride 24 elements (48 bytes:
// the eight rows of a 128-byte LDS access fall on disjoint banks).  Each
// thread stages 8 consecutive k of one row of A and of two rows of B; two
// register sets keep the loads of slab s+2 in flight while slab s is
// multiplied.  Same numerical contract as lx_unit_gemm_bx6.
#define LX_BXW_KS 16
#define LX_BXW_STR 24
struct lx_bxw_smem {
    static constexpr int bytes = 3 * 384 * LX_BXW_STR * 2;
};

template <bool H16 = false>
__device__ __forceinline__ void lx_unit_gemm_bx6w(float *__restrict__ C, long long ldc, const float *__restrict__ A,
                                                  long long lda, const float *__restrict__ B, long long ldb, int M,
                                                  int N, int K, bool lower, int gdiag, char *smem)
{
    typedef lx_split<H16> S;
    typedef typename S::e E;
    typedef typename S::v V;
    constexpr int KS = LX_BXW_KS, STR = LX_BXW_STR;
    E(*sP)[384][STR] = reinterpret_cast<E(*)[384][STR]>(smem);
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave();
    const int wr = (w >> 1) * 64, wc = (w & 1) * 128;
    const bool active = wr < M && wc < N && !(lower && wr + 63 + gdiag < wc);
    const int nslab = (K + KS - 1) / KS;
    // staging: thread -> row srow of A, rows srow and srow + 128 of B, k from sk
    // (rows clamped, never stored; k past K zero)
    const int srow = tid & 127, sk = (tid >> 7) * 8;
    const float *pa = A + (srow < M ? srow : M - 1);
    const float *pb0 = B + (srow < N ? srow : N - 1);
    const float *pb1 = B + (srow + 128 < N ? srow + 128 : N - 1);
    float xa[2][8], xb0[2][8], xb1[2][8];
    const bool kfull = K % KS == 0;  // uniform
    auto gload = [&](int set, int k0) {
        if (kfull) {
            const long long kk = k0 + sk;
            const float *qa = pa + kk * lda, *q0 = pb0 + kk * ldb, *q1 = pb1 + kk * ldb;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                xa[set][j] = qa[(long long)j * lda];
                xb0[set][j] = q0[(long long)j * ldb];
                xb1[set][j] = q1[(long long)j * ldb];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + sk + j;
                const bool kv = k < K;
                const long long kc = kv ? k : K - 1;
                const float va = pa[kc * lda], v0 = pb0[kc * ldb], v1 = pb1[kc * ldb];
                xa[set][j] = kv ? va : 0.f;
                xb0[set][j] = kv ? v0 : 0.f;
                xb1[set][j] = kv ? v1 : 0.f;
            }
        }
    };
    auto lstore = [&](int set) {
        V pl[3];
        S::split(xa[set], pl);
#pragma unroll
        for (int q = 0; q < S::np; ++q) *(V *)&sP[q][srow][sk] = pl[q];
        S::split(xb0[set], pl);
#pragma unroll
        for (int q = 0; q < S::np; ++q) *(V *)&sP[q][128 + srow][sk] = pl[q];
        S::split(xb1[set], pl);
#pragma unroll
        for (int q = 0; q < S::np; ++q) *(V *)&sP[q][256 + srow][sk] = pl[q];
    };
    lx_f16v acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;
    const int lr = lane & 31, lk = lane >> 5;
    auto compute = [&]() {
        if (!active) return;
        V a[2][S::np];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int p = 0; p < S::np; ++p) a[i][p] = *(const V *)&sP[p][wr + i * 32 + lr][8 * lk];
        // two column tiles at a time: four independent accumulators between
        // consecutive MFMAs on one of them
#pragma unroll
        for (int jp = 0; jp < 4; jp += 2) {
            V b[2][S::np];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int p = 0; p < S::np; ++p) b[j][p] = *(const V *)&sP[p][128 + wc + (jp + j) * 32 + lr][8 * lk];
#pragma unroll
            for (int q = 0; q < S::nq; ++q)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][jp + j] = S::mma(a[i][S::pa(q)], b[j][S::pb(q)], acc[i][jp + j]);
        }
    };
    gload(0, 0);
    // the previous block of the same unit may still be reading the planes
    __syncthreads();
    lstore(0);
    __syncthreads();
    if (nslab > 1) gload(0, KS);
    if (nslab > 2) gload(1, 2 * KS);
    // unrolled by two so that the register sets have fixed names: on entry
    // slab sl is in LDS, slab sl+1 in set 0 and slab sl+2 in set 1
    for (int sl = 0; sl < nslab; sl += 2) {
        compute();  // slab sl
        if (sl + 1 < nslab) {
            __syncthreads();
            lstore(0);  // slab sl+1
            __syncthreads();
            if (sl + 3 < nslab) gload(0, (sl + 3) * KS);
            compute();  // slab sl+1
            if (sl + 2 < nslab) {
                __syncthreads();
                lstore(1);  // slab sl+2
                __syncthreads();
                if (sl + 4 < nslab) gload(1, (sl + 4) * KS);
            }
        }
    }
    if (active) {
        // C -= acc, two tiles at a time, their loads before their stores
        const int lane2 = lx_lane(), lr2 = lane2 & 31, lk2 = lane2 >> 5;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jp = 0; jp < 4; jp += 2) {
                float v[2][16];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int c = wc + (jp + j) * 32 + lr2, cc = c < N ? c : N - 1;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int r = wr + i * 32 + 8 * (e / 4) + 4 * lk2 + (e % 4), rr = r < M ? r : M - 1;
                        v[j][e] = C[(long long)cc * ldc + rr];
                    }
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int c = wc + (jp + j) * 32 + lr2;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int r = wr + i * 32 + 8 * (e / 4) + 4 * lk2 + (e % 4);
                        if (r < M && c < N && (!lower || r + gdiag >= c))
                            C[(long long)c * ldc + r] = v[j][e] - S::unscale(acc[i][jp + j][e]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
    }
}

// ---------------------------------------------------------------------------
// The split-FP16 trailing update from pre-split operands (lx_sys::planes):
// C[M x N] -= A[M x K] B[N x K]^T with M <= 128, N <= 256, K a multiple of 16
// and the operands' columns starting on a 16-column slab boundary.  The hi and
// lo FP16 planes of 2^12 x were written once, by the TRSM unit that produced
// x (lx_trsm_write_planes), in slabs of 16 columns:
//   plane p, slab g, row r -> halves [((p nsig + g) ldp + r) 16, + 16)
// so 32 rows of one slab of one plane are a contiguous KB whose 64 chunks of
// 16 bytes are, in order, the v_mfma_f32_32x32x16_f16 operand of lanes 0..63
// (lane l: row l % 32, columns 8 (l / 32) .. + 7).  A slab of the block's
// operands (A: 4 tiles of 32 rows, B: 8, two planes each: 24 blocks of 1 KB)
// is staged into LDS in that order by 16-byte loads, six per thread, and each
// wave reads an operand tile back with one 16-byte LDS read per lane at 16 l:
// no split and no conversion in the loop, and no LDS bank conflicts.  Two LDS
// buffers and two register sets: slab s is multiplied from one buffer while
// slab s+1 is stored to the other and slabs s+2 and s+3 are in flight.  The
// products (lo hi, hi lo, hi hi), their order and the FP32 accumulation are
// those of lx_unit_gemm_bx6w<true>, from the same FP32 values of L, so the
// result is the same bit for bit.
#define LX_P16_BLK 24
struct lx_p16_smem {
    static constexpr int bytes = 2 * LX_P16_BLK * 1024;
};

__device__ __forceinline__ void lx_unit_gemm_p16(float *__restrict__ C, long long ldc, const _Float16 *__restrict__ P,
                                                 long long ldp, int nsig, int arow, int brow, int g0, int M, int N,
                                                 int K, bool lower, int gdiag, char *smem)
{
    typedef lx_split<true> S;
    typedef S::v V;
    const int tid = lx_tid(), lane = tid & 63, w = __builtin_amdgcn_readfirstlane(lx_wave());
    const int wr = (w >> 1) * 64, wc = (w & 1) * 128;
    const bool active = wr < M && wc < N && !(lower && wr + 63 + gdiag < wc);
    const int nslab = K / 16;
    // chunk m of this thread: operand block 4 m + w (blocks 2 t + p: A tile t
    // of 4, then B tile t of 8, plane p), lane `lane` of it.  The block's
    // address is wave-uniform (scalar registers); the lane's offset inside it,
    // row lane % 32 and half lane / 32, is the same for every block
    const long long sstride = ldp * 16;
    const int loff = (lane & 31) * 16 + 8 * (lane >> 5);
    auto blk_base = [&](int m) {
        const int blk = 4 * m + w, isb = blk >= 8, t = (isb ? blk - 8 : blk) >> 1, p = blk & 1;
        return P + ((long long)(p * nsig + g0) * ldp + (isb ? brow : arow) + 32 * t) * 16;
    };
    V xs[2][6];
    auto gload = [&](int set, int sl) {
#pragma unroll
        for (int m = 0; m < 6; ++m) xs[set][m] = *(const V *)(blk_base(m) + sl * sstride + loff);
    };
    auto lstore = [&](int buf, int set) {
        char *b = smem + buf * LX_P16_BLK * 1024 + 16 * lane;
#pragma unroll
        for (int m = 0; m < 6; ++m) *(V *)(b + (4 * m + w) * 1024) = xs[set][m];
    };
    lx_f16v acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;
    auto compute = [&](int buf) {
        if (!active) return;
        const char *b0 = smem + buf * LX_P16_BLK * 1024 + 16 * lane;
        V a[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int p = 0; p < 2; ++p) a[i][p] = *(const V *)(b0 + (2 * (wr / 32 + i) + p) * 1024);
        // two column tiles at a time: four independent accumulators between
        // consecutive MFMAs on one of them
#pragma unroll
        for (int jp = 0; jp < 4; jp += 2) {
            V b[2][2];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int p = 0; p < 2; ++p) b[j][p] = *(const V *)(b0 + (8 + 2 * (wc / 32 + jp + j) + p) * 1024);
#pragma unroll
            for (int q = 0; q < S::nq; ++q)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][jp + j] = S::mma(a[i][S::pa(q)], b[j][S::pb(q)], acc[i][jp + j]);
        }
    };
    gload(0, 0);
    if (nslab > 1) gload(1, 1);
    // the previous block of the same unit may still be reading the buffers
    __syncthreads();
    lstore(0, 0);
    __syncthreads();
    if (nslab > 2) gload(0, 2);
    // on entry: buffer 0 holds slab sl, set 1 slab sl+1, set 0 slab sl+2
    for (int sl = 0; sl < nslab; sl += 2) {
        compute(0);
        if (sl + 1 < nslab) {
            lstore(1, 1);
            __syncthreads();  // slab sl+1 visible; every wave is past compute(0)
            if (sl + 3 < nslab) gload(1, sl + 3);
            compute(1);
            if (sl + 2 < nslab) {
                lstore(0, 0);
                __syncthreads();  // slab sl+2 visible; every wave is past compute(1)
                if (sl + 4 < nslab) gload(0, sl + 4);
            }
        }
    }
    if (active) {
        // C -= acc, two tiles at a time, their loads before their stores
        const int lane2 = lx_lane(), lr2 = lane2 & 31, lk2 = lane2 >> 5;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jp = 0; jp < 4; jp += 2) {
                float v[2][16];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int c = wc + (jp + j) * 32 + lr2, cc = c < N ? c : N - 1;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int r = wr + i * 32 + 8 * (e / 4) + 4 * lk2 + (e % 4), rr = r < M ? r : M - 1;
                        v[j][e] = C[(long long)cc * ldc + rr];
                    }
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int c = wc + (jp + j) * 32 + lr2;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int r = wr + i * 32 + 8 * (e / 4) + 4 * lk2 + (e % 4);
                        if (r < M && c < N && (!lower || r + gdiag >= c))
                            C[(long long)c * ldc + r] = v[j][e] - S::unscale(acc[i][jp + j][e]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
    }
}

// ---------------------------------------------------------------------------
// Wave-level C_acc += A * B^T with A from global memory (rows r0.. of a
// column-major panel, clamped past M) and B^T from LDS: element (k, c) of B^T
// is sB[c * ldsb + k] (row c of a row-major 64 x K block).  16 rows x 16*FN
// columns, K a multiple of 4 (columns >= K must hold zeros in sB).
template <typename T, int FN>
__device__ __forceinline__ void lx_wave_mm_ldsb(typename lx_mfma<T>::acc (&acc)[1][FN], const T *__restrict__ A,
                                                long long lda, int r0, int M, const T *sB, int ldsb, int K)
{
    const int lane = lx_lane(), lr = lane & 15, lk = lane >> 4;
    const int r = r0 + lr < M ? r0 + lr : M - 1;
    const T *pa = A + r;
    if (K == 64) {  // uniform: the POTRF panel; all 16 loads in flight at once
        T a[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) a[s] = pa[(long long)(4 * s + lk) * lda];
#pragma unroll
        for (int s = 0; s < 16; ++s)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[0][j] = lx_mfma<T>::mac(a[s], sB[(j * 16 + lr) * ldsb + 4 * s + lk], acc[0][j]);
        return;
    }
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 4) {
        const T a = pa[(long long)(k0 + lk) * lda];
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[0][j] = lx_mfma<T>::mac(a, sB[(j * 16 + lr) * ldsb + k0 + lk], acc[0][j]);
    }
}

// Value of v in lane l (l wave-uniform), for every lane: v_readlane, no LDS.
__device__ __forceinline__ float lx_readlane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lx_readlane(double v, int l)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// ---------------------------------------------------------------------------
// Z = L^-1 for a 64 x 64 lower-triangular L held in LDS (sL[r][c], row
// stride 65; the upper blocks of sL are used as scratch and lost), by blocks:
// each wave inverts one 16 x 16 diagonal block (lane = column), then
// Z_IJ = -Z_II sum_{K=J}^{I-1} L_IK Z_KJ on the matrix cores for I = 1..3.
// sZ must be zero on entry and sL/sZ visible to the whole workgroup; all four
// waves must call it (it contains workgroup barriers); sZ is complete and
// visible on return.
// 1/sqrt(x) for x > 0 to about 1 ulp: the hardware estimate, then Newton
// steps r <- r + r (1 - x r^2) / 2 (two for FP64, one for FP32).  A short
// dependency chain (the Cholesky pivots are sequential); x <= 0 gives NaN or
// infinity, as 1/sqrt(x) does.
template <typename T> __device__ __forceinline__ T lx_rsqrt(T x)
{
    if constexpr (sizeof(T) == 8) {
        double r = __builtin_amdgcn_rsq(x);
#pragma unroll
        for (int it = 0; it < 2; ++it) r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
        return r;
    } else {
        float r = __builtin_amdgcn_rsqf(x);
        return fmaf(0.5f * r, fmaf(-x * r, r, 1.0f), r);
    }
}

template <typename T>
__device__ __forceinline__ void lx_tri_inv64(T (*sL)[65], T (*sZ)[65])
{
    const int lane = lx_lane(), w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    {
        const int b0 = 16 * w, c = lr;
        T z[16], id[16];
        // the reciprocals of the diagonal first: no division on the chain
        // of the substitution
#pragma unroll
        for (int r = 0; r < 16; ++r) id[r] = T(1) / sL[b0 + r][b0 + r];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            T s = r == c ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < r; ++k) s -= sL[b0 + r][b0 + k] * z[k];
            z[r] = r >= c ? s * id[r] : T(0);
        }
        if (lane < 16) {
#pragma unroll
            for (int r = 0; r < 16; ++r) sZ[b0 + r][b0 + c] = z[r];
        }
    }
    __syncthreads();
    // the partial sum S of block (I, J) goes through sL's unused upper block (J, I)
    for (int I = 1; I < 4; ++I) {
        if (w < I) {  // wave-uniform
            const int J = w;
            typename lx_mfma<T>::acc S[1][1], R[1][1];
            lx_acc_zero<T, 1, 1>(S);
            lx_acc_zero<T, 1, 1>(R);
            for (int k0 = 0; k0 < 16 * (I - J); k0 += 4)
                S[0][0] = lx_mfma<T>::mac(sL[16 * I + lr][16 * J + k0 + lk], sZ[16 * J + k0 + lk][16 * J + lr], S[0][0]);
#pragma unroll
            for (int i = 0; i < 4; ++i) sL[16 * J + lx_mfma<T>::row(lk, i)][16 * I + lr] = S[0][0][i];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k0 = 0; k0 < 16; k0 += 4)
                R[0][0] = lx_mfma<T>::mac(sZ[16 * I + lr][16 * I + k0 + lk], sL[16 * J + k0 + lk][16 * I + lr], R[0][0]);
#pragma unroll
            for (int i = 0; i < 4; ++i) sZ[16 * I + lx_mfma<T>::row(lk, i)][16 * J + lr] = -R[0][0][i];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The TRSM unit's update of one 64-column block:
//   acc (wave w: rows 16w..16w+15, the block's 64 columns) += X[:, <K] Lq[:, <K]^T
// with X the unit's R <= 64 rows (ld ldx) and Lq the block's qb <= 64 rows of
// L (ld ldl), K a multiple of 64.  Both operands go through LDS in K-slabs
// shared by the four waves ([k][row], double buffered; 16-byte loads of slab
// s+1 in flight while slab s is multiplied), so each element of L is loaded
// once per workgroup instead of once per wave, and the slab loads overlap.
// Ragged blocks (R or qb < 64) take element loads with the rows clamped (the
// clamped rows are never stored).  LDS: 4 * KS * STR * sizeof(T) bytes
// (lx_trsm_upd_smem<T>::bytes), any 16-byte aligned arena.
template <typename T> struct lx_trsm_upd_smem {
    static constexpr int STR = sizeof(T) == 8 ? 68 : 80;
    static constexpr int bytes = 4 * lx_lgemm_cfg<T>::KS * STR * (int)sizeof(T);
};
template <typename T>
__device__ __forceinline__ void lx_trsm_update_lds(typename lx_mfma<T>::acc (&acc)[1][4], const T *__restrict__ X,
                                                   long long ldx, int R, const T *__restrict__ Lq, long long ldl,
                                                   int qb, int K, char *smem)
{
    constexpr int KS = lx_lgemm_cfg<T>::KS, VN = lx_vec16<T>::n, STR = lx_trsm_upd_smem<T>::STR;
    typedef typename lx_vec16<T>::type V;
    constexpr int VPC = 64 / VN;                        // vectors per slab column
    constexpr int NV = 64 * KS / VN / LX_WG_THREADS;    // vectors per thread per operand and slab
    T(*sX)[KS][STR] = reinterpret_cast<T(*)[KS][STR]>(smem);
    T(*sL)[KS][STR] = reinterpret_cast<T(*)[KS][STR]>(smem + 2 * KS * STR * sizeof(T));
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    const int nslab = K / KS;
    auto compute = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            const T a = sX[buf][kk + lk][16 * w + lr];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[0][j] = lx_mfma<T>::mac(a, sL[buf][kk + lk][16 * j + lr], acc[0][j]);
        }
    };
    if (R == 64 && qb == 64) {  // uniform
        V rx[NV], rl[NV];
        auto gload = [&](int k0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int q = tid + i * LX_WG_THREADS, kk = q / VPC, r = (q % VPC) * VN;
                rx[i] = *(const V *)(X + (long long)(k0 + kk) * ldx + r);
                rl[i] = *(const V *)(Lq + (long long)(k0 + kk) * ldl + r);
            }
        };
        auto lstore = [&](int buf) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int q = tid + i * LX_WG_THREADS, kk = q / VPC, r = (q % VPC) * VN;
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    sX[buf][kk][r + e] = rx[i][e];
                    sL[buf][kk][r + e] = rl[i][e];
                }
            }
        };
        gload(0);
        lstore(0);
        __syncthreads();
        for (int sl = 0; sl < nslab; ++sl) {
            const int buf = sl & 1;
            if (sl + 1 < nslab) gload((sl + 1) * KS);
            compute(buf);
            if (sl + 1 < nslab) lstore(buf ^ 1);
            __syncthreads();
        }
    } else {
        for (int sl = 0; sl < nslab; ++sl) {
            const int k0 = sl * KS;
#pragma unroll 4
            for (int q = 0; q < 64 * KS / LX_WG_THREADS; ++q) {
                const int e = tid + q * LX_WG_THREADS, r = e % 64, kk = e / 64;
                sX[0][kk][r] = X[(long long)(k0 + kk) * ldx + (r < R ? r : R - 1)];
                sL[0][kk][r] = Lq[(long long)(k0 + kk) * ldl + (r < qb ? r : qb - 1)];
            }
            __syncthreads();
            compute(0);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal-tile Cholesky (reference task_dpotrf -> dpotrf_("L"), tasks_potrf.c
// :20-24), blocked right-looking inside the tile with 64-wide panels.  For
// each panel the 64 x 64 diagonal block D sits in LDS (rows past the tile are
// an identity pad) and:
//   1) D is factored left-looking in four 16-column blocks: the matrix cores
//      subtract the earlier columns (wave w owns rows 16w..16w+15), then
//      wavefront 0 factors the 16 columns with one row per lane, held in
//      registers (v_readlane broadcasts the pivots and the scaled column);
//   2) Z = D^-1 by blocks (lx_tri_inv64); Z goes to LDS and to the workspace
//      W (the panel TRSM units of this step read it there);
//   3) the rows below D become A_panel Z^T on the matrix cores (each wave owns
//      16-row chunks, in place);
//   4) the lower trailing part is updated with K = 64 on the matrix cores.
// Returns through *info (same value in every thread) the 1-based local column
// of the first non-positive pivot, as LAPACK's info; the factorisation keeps
// going past it (as LAPACK, results there are not meaningful).
// LDS: max(2 * 64 * 65 * sizeof(T) + 16, lx_lgemm_smem<T>::bytes) bytes.
// Nc < M factors only the leading Nc columns of the M-row matrix (a tall
// panel): the same steps, with the trailing updates restricted to the
// columns left of Nc (the caller updates the rest with the whole panel).
template <typename T>
__device__ __forceinline__ void lx_unit_potrf(T *__restrict__ A, long long lda, int M, T *__restrict__ W,
                                              long long ldw, int *info_local, char *smem, int Nc = -1)
{
    constexpr int PB = 64, LD = 65;
    T(*sD)[LD] = reinterpret_cast<T(*)[LD]>(smem);
    T(*sZ)[LD] = reinterpret_cast<T(*)[LD]>(smem + sizeof(T) * PB * LD);
    int *s_bad = reinterpret_cast<int *>(smem + 2 * sizeof(T) * PB * LD);
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    const int NC = (Nc < 0 || Nc > M) ? M : Nc;
    int first_bad = 0;
    for (int p = 0; p < NC; p += PB) {
        const int pb = M - p < PB ? M - p : PB;
        // load D (full square, identity pad past pb); clear Z.  The loads are
        // unconditional, their indices clamped into the block (a load under a
        // per-lane condition is a branch, and each one then waits for its own
        // data before the next issues); a constant trip count, unrolled, so
        // several loads are in flight per thread
#pragma unroll 4
        for (int q = 0; q < PB * PB / LX_WG_THREADS; ++q) {
            const int e = tid + q * LX_WG_THREADS, r = e % PB, c = e / PB;
            const T v = A[(long long)(p + (c < pb ? c : pb - 1)) * lda + p + (r < pb ? r : pb - 1)];
            sD[r][c] = (r < pb && c < pb) ? v : (r == c ? T(1) : T(0));
            sZ[r][c] = T(0);
        }
        __syncthreads();
        // 1) factor D, 16 columns at a time
        for (int J = 0; J < 4; ++J) {
            const int j0 = 16 * J;
            if (J > 0 && w >= J) {  // wave-uniform: D[rows of wave w, j0:j0+16] -= L[., 0:j0] L[j0:j0+16, 0:j0]^T
                typename lx_mfma<T>::acc acc[1][1];
                lx_acc_zero<T, 1, 1>(acc);
                for (int k0 = 0; k0 < j0; k0 += 4)
                    acc[0][0] = lx_mfma<T>::mac(sD[16 * w + lr][k0 + lk], sD[j0 + lr][k0 + lk], acc[0][0]);
#pragma unroll
                for (int i = 0; i < 4; ++i) sD[16 * w + lx_mfma<T>::row(lk, i)][j0 + lr] -= acc[0][0][i];
            }
            __syncthreads();
            if (w == 0) {
                T a[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) a[j] = sD[lane][j0 + j];
                // lanes above the diagonal of a column are the upper triangle,
                // never read as data (the inverse below writes its scratch
                // there before reading it; only the lower part is stored), so
                // every lane takes the same arithmetic, without selects
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const int cc = j0 + c;
                    const T piv = lx_readlane(a[c], cc);
                    if (cc < pb && !(piv > T(0)) && first_bad == 0) first_bad = p + cc + 1;  // uniform
                    const T rd = lx_rsqrt<T>(piv);
                    a[c] = lane == cc ? piv * rd : a[c] * rd;
#pragma unroll
                    for (int c2 = c + 1; c2 < 16; ++c2) a[c2] -= a[c] * lx_readlane(a[c], j0 + c2);
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) sD[lane][j0 + j] = a[j];
                s_bad[0] = first_bad;  // every lane of wave 0 stores the same value
            }
            __syncthreads();
        }
        first_bad = s_bad[0];
        // 2) Z = D^-1 (uses D's upper blocks as scratch)
        lx_tri_inv64<T>(sD, sZ);
        // Z -> W (the full 64 x 64 block; W holds n + 64 rows, so p + 63 stays
        // inside), L_pp (lower) -> A
        for (int e = tid; e < PB * PB; e += LX_WG_THREADS) {
            const int r = e % PB, c = e / PB;
            W[(long long)c * ldw + p + r] = sZ[r][c];
            if (r < pb && c < pb && r >= c) A[(long long)(p + c) * lda + p + r] = sD[r][c];
        }
        // 3) panel rows below: P := P * Z^T, 16-row chunks per wave, in place
        const int r_beg = p + pb, nrow = M - r_beg;
        const int nchunk = (nrow + 15) / 16;
        T *P = A + (long long)p * lda;
        for (int ch = w; ch < nchunk; ch += 4) {
            typename lx_mfma<T>::acc acc[1][4];
            lx_acc_zero<T, 1, 4>(acc);
            lx_wave_mm_ldsb<T, 4>(acc, P, lda, r_beg + ch * 16, M, &sZ[0][0], LD, PB);
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = r_beg + ch * 16 + lx_mfma<T>::row(lk, e), c = j * 16 + lr;
                    if (r < M && c < pb) P[(long long)c * lda + r] = acc[0][j][e];
                }
        }
        __syncthreads();
        // 4) trailing update of A[r_beg:M, r_beg:M] (lower) with the panel, K = pb,
        // in staged 128 x 128 blocks (D and Z are no longer needed: the blocks
        // take the whole LDS arena); a tall panel only updates its own columns
        if (NC < M) {  // uniform
            const int ncol = NC - r_beg;
            if (ncol > 0) {
                const T *Pp = A + (long long)p * lda + r_beg;
                T *Ct = A + (long long)r_beg * lda + r_beg;
                constexpr int BS = 128;
                for (int cc = 0; cc < ncol; cc += BS)
                    for (int rr = cc; rr < nrow; rr += BS)
                        lx_unit_gemm_lds<T>(Ct + (long long)cc * lda + rr, lda, Pp + rr, lda, Pp + cc, lda,
                                            nrow - rr < BS ? nrow - rr : BS, ncol - cc < BS ? ncol - cc : BS, pb,
                                            rr == cc, rr - cc, smem);
            }
        } else if (nrow > 0 && nrow <= 64 && pb == 64) {  // uniform: the last panel of a 128-row tile
            // one 64 x 64 block through the shared LDS slabs of the TRSM update,
            // then C -= acc on and below the diagonal
            const T *Pp = A + (long long)p * lda + r_beg;
            T *Ct = A + (long long)r_beg * lda + r_beg;
            typename lx_mfma<T>::acc acc[1][4];
            lx_acc_zero<T, 1, 4>(acc);
            lx_trsm_update_lds<T>(acc, Pp, lda, nrow, Pp, lda, nrow, PB, smem);
            // every load of C first (indices clamped into the block), then the
            // masked stores: a read-modify-write under the mask would wait a
            // memory round trip per element
            T cv[4][4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * w + lx_mfma<T>::row(lk, e), c = j * 16 + lr;
                    cv[j][e] = Ct[(long long)(c < nrow ? c : nrow - 1) * lda + (r < nrow ? r : nrow - 1)];
                }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * w + lx_mfma<T>::row(lk, e), c = j * 16 + lr;
                    if (r < nrow && c <= r) Ct[(long long)c * lda + r] = cv[j][e] - acc[0][j][e];
                }
        } else if (nrow > 0) {
            constexpr int BS = 128;
            const T *Pp = A + (long long)p * lda + r_beg;
            T *Ct = A + (long long)r_beg * lda + r_beg;
            const int nb = (nrow + BS - 1) / BS;
            for (int bj = 0; bj < nb; ++bj)
                for (int bi = bj; bi < nb; ++bi) {
                    const int rr = bi * BS, cc = bj * BS;
                    lx_unit_gemm_lds<T>(Ct + (long long)cc * lda + rr, lda, Pp + rr, lda, Pp + cc, lda,
                                        nrow - rr < BS ? nrow - rr : BS, nrow - cc < BS ? nrow - cc : BS, pb, bi == bj,
                                        rr - cc, smem);
                }
        }
        __syncthreads();
    }
    *info_local = first_bad;
}

// ---------------------------------------------------------------------------
// Panel TRSM unit (reference task_dtrsm(R, L, T, N), chol_rlmain.c:37 ->
// tasks_trsm.c:22-25):  X[R x N] = B[R x N] * L^{-T}, R <= 64, rows
// independent.  Wavefront w owns rows 16w..16w+15 and the workgroup walks the
// 64-column blocks q left to right:
//   Y = B_q - X_{<q} L[q, <q]^T   (matrix cores, K = 64 q, operands staged
//                                   through LDS by lx_trsm_update_lds)
//   X_q = Y Z_q^T                  (Z_q = inverse of the diagonal 64 x 64
//                                   block, written to W by the POTRF unit and
//                                   staged in LDS once per block)
// Every wave runs every block (waves past R compute clamped rows and store
// nothing), so the barriers per block match in all four.
// LDS: 64 * 65 * sizeof(T) + max(4 * 16 * 65 * sizeof(T), lx_trsm_upd_smem<T>::bytes).
// P (FP32 factor with split-FP16 updates, else null): the plane buffer of
// lx_unit_gemm_p16; the unit then also writes hi and lo of 2^12 X there,
// X being rows grow.. and columns gcol.. of L (gcol a multiple of 16).
template <typename T>
__device__ __forceinline__ void lx_unit_trsm_rltn(T *__restrict__ Bm, long long ldb, const T *__restrict__ L,
                                                  long long ldl, const T *__restrict__ W, long long ldw, int R, int N,
                                                  char *smem, _Float16 *__restrict__ P = nullptr, long long ldp = 0,
                                                  int nsig = 0, int grow = 0, int gcol = 0)
{
    constexpr int QB = 64, LD = 65;
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    T(*sZq)[LD] = reinterpret_cast<T(*)[LD]>(smem);  // sZq[c][k] = Z_q[c][k]
    // the update's slabs, then (after a barrier) sY, both past sZq
    char *sU = smem + sizeof(T) * QB * LD;
    T(*sY)[LD] = reinterpret_cast<T(*)[LD]>(sU + sizeof(T) * 16 * LD * w);
    const int r0 = w * 16;
    const bool active = r0 < R;  // wave-uniform
    for (int q = 0; q < N; q += QB) {
        const int qb = N - q < QB ? N - q : QB;
        // Z_q (column-major in W: Z[r][c] at W[c * ldw + q + r]) into registers
        // now, into LDS after the update: its loads overlap the update
        T zr[QB * QB / LX_WG_THREADS];
#pragma unroll
        for (int q2 = 0; q2 < QB * QB / LX_WG_THREADS; ++q2) {
            const int e = tid + q2 * LX_WG_THREADS, r = e % QB, c = e / QB;
            zr[q2] = W[(long long)c * ldw + q + r];
        }
        // this block's right-hand sides, loaded before the long update
        T bq[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rl = lx_mfma<T>::row(lk, e), c = j * 16 + lr;
                const int rg = r0 + rl < R ? r0 + rl : R - 1;
                bq[j][e] = Bm[(long long)(q + (c < qb ? c : qb - 1)) * ldb + rg];
            }
        typename lx_mfma<T>::acc acc[1][4];
        lx_acc_zero<T, 1, 4>(acc);
        // X_{<q} L[q, <q]^T through LDS (every wave takes part in the staging)
        if (q > 0) lx_trsm_update_lds<T>(acc, Bm, ldb, R, L + q, ldl, qb, q, sU);  // ends with a barrier
        // Y = B_q - acc -> LDS (row-major, zero past the block); Z_q -> LDS
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rl = lx_mfma<T>::row(lk, e), c = j * 16 + lr;
                sY[rl][c] = c < qb ? bq[j][e] - acc[0][j][e] : T(0);
            }
#pragma unroll
        for (int q2 = 0; q2 < QB * QB / LX_WG_THREADS; ++q2) {
            const int e = tid + q2 * LX_WG_THREADS, r = e % QB, c = e / QB;
            sZq[r][c] = zr[q2];
        }
        __syncthreads();  // sZq staged, sY written
        // X_q = Y Z_q^T
        typename lx_mfma<T>::acc x[1][4];
        lx_acc_zero<T, 1, 4>(x);
#pragma unroll 2
        for (int k0 = 0; k0 < QB; k0 += 4) {
            const T a = sY[lr][k0 + lk];
#pragma unroll
            for (int j = 0; j < 4; ++j) x[0][j] = lx_mfma<T>::mac(a, sZq[j * 16 + lr][k0 + lk], x[0][j]);
        }
        if (active) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = r0 + lx_mfma<T>::row(lk, e), c = j * 16 + lr;
                    if (r < R && c < qb) Bm[(long long)(q + c) * ldb + r] = x[0][j][e];
                }
        }
        if constexpr (sizeof(T) == 4) {
            if (P != nullptr) {  // uniform
                // this wave's 16 rows of X_q, row-major through its own sY, then
                // 8 columns of one row per lane and chunk: split into hi and lo
                // of 2^12 x (as lx_unit_gemm_bx6w<true> splits the same values)
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) sY[lx_mfma<T>::row(lk, e)][j * 16 + lr] = x[0][j][e];
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int ch = lane + 64 * h, rl = ch >> 3, c = (ch & 7) * 8;
                    if (active && r0 + rl < R && c + 8 <= qb) {
                        float xv[8];
#pragma unroll
                        for (int e = 0; e < 8; ++e) xv[e] = (float)sY[rl][c + e];
                        lx_split<true>::v pl[3];
                        lx_split<true>::split(xv, pl);
                        const int gc = gcol + q + c;
                        _Float16 *dst = P + ((long long)(gc >> 4) * ldp + grow + r0 + rl) * 16 + 8 * ((gc >> 3) & 1);
                        *(lx_split<true>::v *)dst = pl[0];
                        *(lx_split<true>::v *)(dst + (long long)nsig * ldp * 16) = pl[1];
                    }
                }
            }
        }
        __syncthreads();  // sZq and sY are rewritten by the next block
    }
}

// ---------------------------------------------------------------------------
// Right-hand-side sweeps (reference bblas_dltrsm lower branches,
// src/blocked_blas/bblas_trsm.h:119-160): X has nrhs columns (ldx).

// X_i[M x nrhs] := L_ii^{-1} X_i  (trans = 0)  or  L_ii^{-T} X_i  (trans = 1).
// Blocked by 64 rows; the 64 x 64 diagonal solve runs on one wavefront per
// right-hand side (lane = row), the off-diagonal update on all 256 threads.
template <typename T>
__device__ __forceinline__ void lx_unit_trsv_block(T *__restrict__ X, long long ldx, const T *__restrict__ L, long long ldl,
                                   int M, int nrhs, int trans, char *smem)
{
    constexpr int NB = 64;
    T(*sD)[NB + 1] = reinterpret_cast<T(*)[NB + 1]>(smem);
    T(*sX)[NB] = reinterpret_cast<T(*)[NB]>(smem + sizeof(T) * NB * (NB + 1));
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave();
    const int nblk = (M + NB - 1) / NB;
    for (int cb = 0; cb < nrhs; cb += 16) {
        const int nc = nrhs - cb < 16 ? nrhs - cb : 16;
        for (int q = 0; q < nblk; ++q) {
            const int blk = trans ? nblk - 1 - q : q;
            const int p = blk * NB;
            const int pb = M - p < NB ? M - p : NB;
            // load diagonal block (as L or L^T so the solve is always "lower, forward" in sD)
            for (int e = tid; e < NB * NB; e += LX_WG_THREADS) {
                int r = e % NB, c = e / NB;
                T v = T(0);
                if (r < pb && c < pb) {
                    if (!trans) v = (r >= c) ? L[(long long)(p + c) * ldl + p + r] : T(0);
                    else v = (c >= r) ? L[(long long)(p + r) * ldl + p + c] : T(0);  // (L^T)[r][c]=L[c][r]; upper
                }
                sD[r][c] = v;
            }
            for (int e = tid; e < 16 * NB; e += LX_WG_THREADS) {
                int r = e % NB, c = e / NB;
                sX[c][r] = (r < pb && c < nc) ? X[(long long)(cb + c) * ldx + p + r] : T(0);
            }
            __syncthreads();
            // each wave solves columns c = w, w+4, ...; lane = row
            for (int c = w; c < nc; c += 4) {
                T s = sX[c][lane];
                if (!trans) {
                    for (int j = 0; j < pb; ++j) {
                        T xj = __shfl(s, j) / sD[j][j];
                        if (lane == j) s = xj;
                        else if (lane > j) s -= sD[lane][j] * xj;
                    }
                } else {
                    // upper-triangular U = L_pp^T: backward substitution
                    for (int j = pb - 1; j >= 0; --j) {
                        T xj = __shfl(s, j) / sD[j][j];
                        if (lane == j) s = xj;
                        else if (lane < j) s -= sD[lane][j] * xj;
                    }
                }
                sX[c][lane] = s;
            }
            __syncthreads();
            for (int e = tid; e < 16 * NB; e += LX_WG_THREADS) {
                int r = e % NB, c = e / NB;
                if (r < pb && c < nc) X[(long long)(cb + c) * ldx + p + r] = sX[c][r];
            }
            // update the remaining rows of this block
            {
                // forward: rows below, X[r] -= L[r, p:p+pb] x_p for r in [p+pb, M)
                // backward: rows above, X[r] -= sum_j L[p+j][r] x_j for r in [0, p)
                // uniform trip count; out-of-range rows repeat a valid row and do not store
                const int rlo = trans ? 0 : p + pb, rhi = trans ? p : M;
                const int nq = (rhi - rlo + LX_WG_THREADS - 1) / LX_WG_THREADS;
                for (int q = 0; q < nq; ++q) {
                    const int rr = rlo + q * LX_WG_THREADS + tid;
                    const bool live = rr < rhi;
                    const int r = live ? rr : rhi - 1;
                    T s[16];
#pragma unroll
                    for (int c = 0; c < 16; ++c) s[c] = T(0);
                    for (int j = 0; j < pb; ++j) {
                        T l = trans ? L[(long long)r * ldl + p + j] : L[(long long)(p + j) * ldl + r];
#pragma unroll
                        for (int c = 0; c < 16; ++c) s[c] += l * sX[c][j];
                    }
                    if (live)
                        for (int c = 0; c < nc; ++c) X[(long long)(cb + c) * ldx + r] -= s[c];
                }
            }
            __syncthreads();
        }
    }
}

// Y[M x nrhs] -= L[M x K] * Z[K x nrhs]   (trans = 0; forward-sweep GEMM)
// Y[M x nrhs] -= L[K x M]^T * Z[K x nrhs] (trans = 1; backward-sweep GEMM)
// Memory-bound: L is read once, coalesced along its columns.
template <typename T>
__device__ __forceinline__ void lx_unit_gemv_rhs(T *__restrict__ Y, long long ldy, const T *__restrict__ L, long long ldl,
                                 const T *__restrict__ Z, long long ldz, int M, int K, int nrhs, int trans,
                                 char *smem)
{
    constexpr int KC = 64;
    // sRed reuses sT's storage once the K loop of a row block is done
    T(*sZ)[KC] = reinterpret_cast<T(*)[KC]>(smem);
    T(*sT)[KC + 1] = reinterpret_cast<T(*)[KC + 1]>(smem + sizeof(T) * 16 * KC);
    T(*sRed)[16][64] = reinterpret_cast<T(*)[16][64]>(smem + sizeof(T) * 16 * KC);
    const int tid = lx_tid();
    for (int cb = 0; cb < nrhs; cb += 16) {
        const int nc = nrhs - cb < 16 ? nrhs - cb : 16;
        if (!trans) {
            // thread: row r = r0 + (tid & 63) handled by 4 k-groups (tid >> 6)
            for (int r0 = 0; r0 < M; r0 += 64) {
                const int r = r0 + (tid & 63), g = tid >> 6;
                T s[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) s[c] = T(0);
                for (int k0 = 0; k0 < K; k0 += KC) {
                    __syncthreads();
                    for (int e = tid; e < 16 * KC; e += LX_WG_THREADS) {
                        int k = e % KC, c = e / KC;
                        sZ[c][k] = (k0 + k < K && c < nc) ? Z[(long long)(cb + c) * ldz + k0 + k] : T(0);
                    }
                    __syncthreads();
                    const int kend = K - k0 < KC ? K - k0 : KC;
                    const int rc = r < M ? r : M - 1;  // rows past M accumulate row M-1, never stored
                    for (int k = g; k < kend; k += 4) {
                        T l = L[(long long)(k0 + k) * ldl + rc];
#pragma unroll
                        for (int c = 0; c < 16; ++c) s[c] += l * sZ[c][k];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int c = 0; c < 16; ++c) sRed[g][c][tid & 63] = s[c];
                __syncthreads();
                for (int c = 0; c < nc; ++c) {
                    const int t = tid & 63;
                    const T v = (sRed[0][c][t] + sRed[1][c][t]) + (sRed[2][c][t] + sRed[3][c][t]);
                    if (tid < 64 && r0 + t < M) Y[(long long)(cb + c) * ldy + r0 + t] -= v;
                }
                __syncthreads();
            }
        } else {
            // output row r (a column of L); stage L[k0:k0+64, r0:r0+64] transposed in LDS
            for (int r0 = 0; r0 < M; r0 += 64) {
                const int rl = tid & 63, g = tid >> 6;
                T s[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) s[c] = T(0);
                for (int k0 = 0; k0 < K; k0 += KC) {
                    __syncthreads();
                    for (int e = tid; e < 16 * KC; e += LX_WG_THREADS) {
                        int k = e % KC, c = e / KC;
                        sZ[c][k] = (k0 + k < K && c < nc) ? Z[(long long)(cb + c) * ldz + k0 + k] : T(0);
                    }
                    for (int e = tid; e < 64 * KC; e += LX_WG_THREADS) {
                        int k = e % KC, rr = e / KC;
                        sT[rr][k] = (k0 + k < K && r0 + rr < M) ? L[(long long)(r0 + rr) * ldl + k0 + k] : T(0);
                    }
                    __syncthreads();
                    for (int k = g; k < KC; k += 4) {
                        T l = sT[rl][k];
#pragma unroll
                        for (int c = 0; c < 16; ++c) s[c] += l * sZ[c][k];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int c = 0; c < 16; ++c) sRed[g][c][rl] = s[c];
                __syncthreads();
                for (int c = 0; c < nc; ++c) {
                    const int t = tid & 63;
                    const T v = (sRed[0][c][t] + sRed[1][c][t]) + (sRed[2][c][t] + sRed[3][c][t]);
                    if (tid < 64 && r0 + t < M) Y[(long long)(cb + c) * ldy + r0 + t] -= v;
                }
                __syncthreads();
            }
        }
    }
}
