// lx_batch.hip -- many independent SPD systems of uneven order, each solved to
// FP64 accuracy by an FP32 Cholesky factor plus FP64 iterative refinement, in
// one persistent launch.
//
// The refinement contract is the single-system one (itref.c:999-1101, restated
// by linompss_dposv_ir_dev): x0 = solve(b); r = b - A x in FP64; repeat
// { berr = ||r||_inf / (||A||_inf ||x||_inf); stop at tol; fall back to the
// FP64 factor on a stall or after max_sweeps; d = solve(r); x += d }.  The
// factor is the reference's tile algorithm applied to one tile
// (task_dpotrf / tasks_potrf.c:20-24 over the whole system), the solves are
// bblas_dltrsm's forward and backward sweeps (bblas_trsm.h:119-160).
//
// Execution model: a workgroup owns one system at a time.  Workgroups take
// systems from a ticket counter in the order the host chose (largest first by
// default, a longest-processing-time schedule), so the launch is a persistent
// work queue with no cross-workgroup waits.  Per system, all 256 threads:
//   1. copy the lower triangle of A to FP32 (workspace slot of the workgroup);
//   2. factor it in place with lx_unit_potrf (64-column panels on the matrix
//      cores, inverses of the 64 x 64 diagonal blocks into W);
//   3. right-looking sweeps over 64-row blocks with the vector in LDS:
//      y_q = Z_q v_q with Z_q the inverse of the diagonal block (from W), then
//      the rows still to be solved subtract their products with y_q;
//   4. FP64 residual r = b - A x over the lower triangle in 64 x 64 tiles
//      (each tile used twice: as A_IJ and, through LDS, as A_IJ^T), which
//      also yields ||A||_inf on the first pass.
// Systems up to LX_BATCH_NMAX rows; one right-hand side each.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lx_device.hpp"

#define LX_BATCH_NMAX 512
#define LX_BATCH_SMEM 73728

struct lx_bsys {
    const double *A;  // n x n, lower triangle read, column-major
    const double *B;  // n
    double *X;        // n (output)
    int64_t lda;
    int n;
    int pad;
};

struct lx_bstat {
    int sweeps;    // corrections applied
    int fallback;  // 1: the FP64 factor produced x
    int info;      // info of the factor that produced x (0 = SPD)
    int pad;
    double berr;   // final normwise backward error
};

// Per-workgroup workspace slot.
struct lx_bslot {
    char *F;      // factor: ld * n elements of the factor's type (room for FP64)
    double *W;    // (ldw) x 64: inverses of the diagonal blocks (room for FP64)
    double *r;    // residual, nmax
    double *v;    // right-hand side / solution of the inner solve, nmax (FP32 or FP64)
    int64_t ldw;
};

__device__ __forceinline__ double lx_b_nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

// Largest of v over the workgroup (NaN wins); every thread gets the result.
__device__ __forceinline__ double lx_b_max(double v, double *red)
{
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = lx_b_nanmax(v, __shfl_xor(v, o));
    if (lane == 0) red[w] = v;
    __syncthreads();
    const double m = lx_b_nanmax(lx_b_nanmax(red[0], red[1]), lx_b_nanmax(red[2], red[3]));
    __syncthreads();
    return m;
}

// L (ld) := lower triangle of A, converted to T, by 64 x 64 tiles: thread
// (row l, columns 16g..16g+15) issues its 16 loads of a tile together.
template <typename T>
__device__ __forceinline__ void lx_b_lower_to(const double *__restrict__ A, int64_t lda, int n, T *__restrict__ L,
                                              int64_t ld)
{
    const int tid = lx_tid(), l = tid & 63, g = tid >> 6;
    const int nb = (n + 63) / 64;
    // the next tile's 16 loads are issued before the current tile is stored
    // (the compiler cannot move loads of A above stores to L itself)
    double an[16];
    auto load_tile = [&](int I, int J) {
        const int row = I * 64 + l;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int col = J * 64 + g * 16 + k;
            an[k] = (row < n && col < n && row >= col) ? A[(int64_t)col * lda + row] : 0.0;
        }
    };
    load_tile(0, 0);
    for (int J = 0; J < nb; ++J)
        for (int I = J; I < nb; ++I) {
            const int row = I * 64 + l;
            double a[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = an[k];
            if (I + 1 < nb) load_tile(I + 1, J);  // uniform
            else if (J + 1 < nb) load_tile(J + 1, J + 1);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int col = J * 64 + g * 16 + k;
                if (row < n && col < n && row >= col) L[(int64_t)col * ld + row] = (T)a[k];
            }
        }
    __syncthreads();
}

// One triangular sweep on v (n entries): trans 0 solves L y = v, trans 1
// solves L^T x = v, with W the inverses Z_q of the 64 x 64 diagonal blocks
// (W[c*ldw + 64q + r] = Z_q[r][c]).  Right-looking over 64-row blocks, v held
// in LDS: y_q = Z_q v_q (Z_q^T for trans 1), then every row r still to be
// solved takes v_r -= sum_c M[r][c] y_q[c], one thread per row with its 64
// loads independent of each other: M[r][c] = L[r][64q + c] (forward, down
// the block's columns) or L[64q + c][r] (backward, along column r).
template <typename T>
__device__ __forceinline__ void lx_b_sweep(const T *__restrict__ L, int64_t ld, const T *__restrict__ W,
                                           int64_t ldw, int n, int trans, T *__restrict__ v, char *smem)
{
    T(*sZ)[65] = reinterpret_cast<T(*)[65]>(smem);
    T(*sred)[64] = reinterpret_cast<T(*)[64]>(smem + sizeof(T) * 64 * 65);
    T *sy = reinterpret_cast<T *>(smem + sizeof(T) * (64 * 65 + 4 * 64));  // y_q, 64
    T *sv = reinterpret_cast<T *>(smem + sizeof(T) * (64 * 65 + 5 * 64));  // v, n
    const int tid = lx_tid(), l = tid & 63, g = tid >> 6;
    const int nb = (n + 63) / 64;
    for (int i = tid; i < n; i += LX_WG_THREADS) sv[i] = v[i];
    for (int qi = 0; qi < nb; ++qi) {
        const int q = trans ? nb - 1 - qi : qi, p = q * 64, pb = n - p < 64 ? n - p : 64;
        // four loads in flight per thread, then their LDS stores (left to
        // itself the scheduler of this kernel issues one load at a time)
#pragma unroll
        for (int q0 = 0; q0 < 64 * 64 / LX_WG_THREADS; q0 += 4) {
            T zv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = tid + (q0 + k) * LX_WG_THREADS;
                zv[k] = W[(int64_t)(e >> 6) * ldw + p + (e & 63)];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = tid + (q0 + k) * LX_WG_THREADS;
                sZ[e & 63][e >> 6] = zv[k];
            }
        }
        __syncthreads();  // sZ, and sv after the previous block's updates
        T z = T(0);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int c = g * 16 + k;
            z += (trans ? sZ[c][l] : sZ[l][c]) * (c < pb ? sv[p + c] : T(0));
        }
        sred[g][l] = z;
        __syncthreads();
        if (g == 0) {
            const T y = (sred[0][l] + sred[1][l]) + (sred[2][l] + sred[3][l]);
            sy[l] = l < pb ? y : T(0);
            if (l < pb) sv[p + l] = y;
        }
        __syncthreads();
        // rows still to be solved: [p + pb, n) forward, [0, p) backward
        const int rlo = trans ? 0 : p + pb, rhi = trans ? p : n;
        for (int r = rlo + tid; r < rhi; r += LX_WG_THREADS) {
            T acc[4] = {T(0), T(0), T(0), T(0)};
            if (!trans) {
                // forward: L[r][p + c] down the block's columns (coalesced over r)
                const T *Lr = L + (int64_t)p * ld + r;
#pragma unroll
                for (int c0 = 0; c0 < 64; c0 += 16) {
                    // unconditional loads (columns clamped into the block, the
                    // clamped ones zeroed after): a load under a condition is
                    // a branch and waits for its data before the next issues
                    T a[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) a[k] = Lr[(int64_t)(c0 + k < pb ? c0 + k : pb - 1) * ld];
#pragma unroll
                    for (int k = 0; k < 16; ++k) acc[k & 3] += (c0 + k < pb ? a[k] : T(0)) * sy[c0 + k];
                    __builtin_amdgcn_sched_barrier(0);  // 16 loads in flight at a time
                }
            } else if (pb == 64) {
                // backward, full block: L^T[r][p + c] = L[p + c][r], the 64 contiguous
                // elements of column r from row p, as 16-byte vectors (ld a multiple
                // of 16 elements, p of 64)
                typedef typename lx_vec16<T>::type V;
                constexpr int VN = lx_vec16<T>::n;
                const T *Lc = L + (int64_t)r * ld + p;
#pragma unroll
                for (int c0 = 0; c0 < 64; c0 += 16) {
                    V a[16 / VN];
#pragma unroll
                    for (int i = 0; i < 16 / VN; ++i) a[i] = *(const V *)(Lc + c0 + i * VN);
#pragma unroll
                    for (int i = 0; i < 16 / VN; ++i)
#pragma unroll
                        for (int e = 0; e < VN; ++e) acc[e & 3] += a[i][e] * sy[c0 + i * VN + e];
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
                // backward, the last (partial) block: rows p + c < n only
                const T *Lc = L + (int64_t)r * ld + p;
#pragma unroll
                for (int c0 = 0; c0 < 64; c0 += 16) {
                    T a[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) a[k] = Lc[c0 + k < pb ? c0 + k : pb - 1];
#pragma unroll
                    for (int k = 0; k < 16; ++k) acc[k & 3] += (c0 + k < pb ? a[k] : T(0)) * sy[c0 + k];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            sv[r] -= (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += LX_WG_THREADS) v[i] = sv[i];
    __syncthreads();
}

// r = b - A x (FP64, A symmetric with its lower triangle stored).  Returns
// max|r|, max|x| and, when want_anorm, max_i sum_j |A_ij| (= ||A||_inf).
__device__ __forceinline__ void lx_b_resid(const double *__restrict__ A, int64_t lda, int n,
                                           const double *__restrict__ x, const double *__restrict__ b,
                                           double *__restrict__ r, bool want_anorm, char *smem, double &rmax,
                                           double &xmax, double &anorm)
{
    double(*sT)[65] = reinterpret_cast<double(*)[65]>(smem);                       // 33280 B
    double(*racc)[LX_BATCH_NMAX] = reinterpret_cast<double(*)[LX_BATCH_NMAX]>(smem + 33280);
    double(*aacc)[LX_BATCH_NMAX] = reinterpret_cast<double(*)[LX_BATCH_NMAX]>(smem + 33280 + 16384);
    double *sx = reinterpret_cast<double *>(smem + 33280 + 2 * 16384);           // LX_BATCH_NMAX
    double *red = sx + LX_BATCH_NMAX;                                             // 4
    const int tid = lx_tid(), l = tid & 63, g = tid >> 6;
    const int nb = (n + 63) / 64;
    for (int i = tid; i < LX_BATCH_NMAX; i += LX_WG_THREADS) {
        sx[i] = i < n ? x[i] : 0.0;
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) {
            racc[gg][i] = 0.0;
            aacc[gg][i] = 0.0;
        }
    }
    __syncthreads();
    // entry (g, i) of racc / aacc is only ever updated by thread (i & 63, g).
    // Tiles in the order (J, I >= J); the 16 loads of the next tile are issued
    // before the current one is used, so a tile's memory round trip overlaps
    // the previous tile's arithmetic and barriers.
    double an[16];
    auto load_tile = [&](int I, int J) {
        const int row = I * 64 + l;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int col = J * 64 + g * 16 + k;
            an[k] = (row < n && col < n && row >= col) ? A[(int64_t)col * lda + row] : 0.0;
        }
    };
    load_tile(0, 0);
    for (int J = 0; J < nb; ++J)
        for (int I = J; I < nb; ++I) {
            const int row = I * 64 + l;
            double a_t[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) a_t[k] = an[k];
            if (I + 1 < nb) load_tile(I + 1, J);  // uniform
            else if (J + 1 < nb) load_tile(J + 1, J + 1);
            double s = 0.0, sa = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int cl = g * 16 + k, col = J * 64 + cl;
                const double a = a_t[k];
                s += a * sx[col < n ? col : 0];
                sa += fabs(a);
                sT[cl][l] = row > col ? a : 0.0;  // the transposed use leaves the diagonal out
            }
            if (row < n) {
                racc[g][row] += s;
                if (want_anorm) aacc[g][row] += sa;
            }
            __syncthreads();
            // (A_IJ^T x_I): thread (column l of the tile, rows 16g..16g+15)
            double t = 0.0, ta = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int rr = g * 16 + k;
                const double a = sT[l][rr];
                t += a * sx[I * 64 + rr < n ? I * 64 + rr : 0];
                ta += fabs(a);
            }
            const int col = J * 64 + l;
            if (col < n) {
                racc[g][col] += t;
                if (want_anorm) aacc[g][col] += ta;
            }
            __syncthreads();
        }
    double rm = 0.0, xm = 0.0, am = 0.0;
    for (int i = tid; i < n; i += LX_WG_THREADS) {
        const double ri = b[i] - ((racc[0][i] + racc[1][i]) + (racc[2][i] + racc[3][i]));
        r[i] = ri;
        rm = lx_b_nanmax(rm, fabs(ri));
        xm = lx_b_nanmax(xm, fabs(sx[i]));
        if (want_anorm) am = lx_b_nanmax(am, (aacc[0][i] + aacc[1][i]) + (aacc[2][i] + aacc[3][i]));
    }
    rmax = lx_b_max(rm, red);
    xmax = lx_b_max(xm, red);
    if (want_anorm) anorm = lx_b_max(am, red);
}

// FP32/FP64 factor of one system, in place: pw = 64 is lx_unit_potrf over the
// whole system (a trailing update per 64-column panel); pw = 128 factors
// 128-column tall panels with it (two 64-column steps each) and updates the
// trailing matrix once per panel with K = 128, so the trailing part is read
// and written half as often.  W and info as for lx_unit_potrf (info: the
// 1-based column of the first non-positive pivot).
template <typename T>
__device__ __forceinline__ void lx_b_potrf(T *__restrict__ F, int64_t ld, int n, T *__restrict__ W, int64_t ldw,
                                           int *info, char *smem, int pw)
{
    if (pw <= 64) {
        lx_unit_potrf<T>(F, ld, n, W, ldw, info, smem);
        return;
    }
    int bad = 0;
    for (int p = 0; p < n; p += pw) {
        const int pc = n - p < pw ? n - p : pw;
        int loc = 0;
        lx_unit_potrf<T>(F + p * ld + p, ld, n - p, W + p, ldw, &loc, smem, pc);
        loc = __builtin_amdgcn_readfirstlane(loc);
        if (loc && !bad) bad = p + loc;
        const int r_beg = p + pc, nrow = n - r_beg;
        if (nrow > 0) {  // uniform: A[r_beg:n, r_beg:n] -= P P^T (lower), P = the panel's rows below
            const T *Pp = F + p * ld + r_beg;
            T *Ct = F + r_beg * ld + r_beg;
            constexpr int BS = 128;
            for (int cc = 0; cc < nrow; cc += BS)
                for (int rr = cc; rr < nrow; rr += BS)
                    lx_unit_gemm_lds<T>(Ct + cc * ld + rr, ld, Pp + rr, ld, Pp + cc, ld, nrow - rr < BS ? nrow - rr : BS,
                                        nrow - cc < BS ? nrow - cc : BS, pc, rr == cc, rr - cc, smem);
            __syncthreads();  // the updated trailing part before the next panel's loads
        }
    }
    *info = bad;
}

// Optional phase profile (prof != null): per workgroup, wall-clock ticks of
// [0] the FP32 copy, [1] the FP32 factor, [2] the first solve, [3] the
// residuals, [4] the correction solves, [5] the FP64 fallback, [6] systems.
#define LX_B_PROF_WORDS 8
struct lx_b_clock {
    long long *acc;  // LDS, LX_B_PROF_WORDS (null: off)
    long long t = 0;
    __device__ void start()
    {
        if (acc) t = wall_clock64();
    }
    __device__ void lap(int q)
    {
        if (!acc) return;
        const long long now = wall_clock64();
        if (lx_tid() == 0) acc[q] += now - t;
        t = now;
    }
};

// Factor + sweeps + refinement of one system by the whole workgroup; pw =
// the FP32 factor's panel width (lx_b_potrf).
__device__ __forceinline__ void lx_b_solve_one(const lx_bsys &S, const lx_bslot &slot, lx_bstat *stat,
                                               int max_sweeps, double tol, double stall, char *smem,
                                               lx_b_clock &clk, int pw)
{
    const int tid = lx_tid();
    const int n = S.n;
    const int64_t ld = ((int64_t)n + 15) & ~(int64_t)15;
    double *red = reinterpret_cast<double *>(smem);
    lx_bstat out = {0, 0, 0, 0, 0.0};
    if (n <= 0) {
        if (tid == 0) *stat = out;
        return;
    }
    // ||b||_inf (the inner solves work on b / ||b|| and r / ||r||)
    double bm = 0.0;
    for (int i = tid; i < n; i += LX_WG_THREADS) bm = lx_b_nanmax(bm, fabs(S.B[i]));
    bm = lx_b_max(bm, red);
    if (bm == 0.0) {  // x = 0 solves it exactly
        for (int i = tid; i < n; i += LX_WG_THREADS) S.X[i] = 0.0;
        __syncthreads();
        if (tid == 0) *stat = out;
        return;
    }
    float *F = reinterpret_cast<float *>(slot.F);
    float *Wf = reinterpret_cast<float *>(slot.W);
    float *vf = reinterpret_cast<float *>(slot.v);
    clk.start();
    lx_b_lower_to<float>(S.A, S.lda, n, F, ld);
    clk.lap(0);
    int info = 0;
    lx_b_potrf<float>(F, ld, n, Wf, slot.ldw, &info, smem, pw);
    info = __builtin_amdgcn_readfirstlane(info);
    clk.lap(1);
    bool fb = info != 0;
    double anorm = 0.0, rmax = 0.0, xmax = 0.0, berr = 0.0;
    if (!fb) {
        for (int i = tid; i < n; i += LX_WG_THREADS) vf[i] = (float)(S.B[i] / bm);
        __syncthreads();
        lx_b_sweep<float>(F, ld, Wf, slot.ldw, n, 0, vf, smem);
        lx_b_sweep<float>(F, ld, Wf, slot.ldw, n, 1, vf, smem);
        for (int i = tid; i < n; i += LX_WG_THREADS) S.X[i] = bm * (double)vf[i];
        __syncthreads();
        clk.lap(2);
        double prev = INFINITY;
        for (int it = 0;; ++it) {
            lx_b_resid(S.A, S.lda, n, S.X, S.B, slot.r, it == 0, smem, rmax, xmax, anorm);
            berr = rmax / (anorm * xmax);
            clk.lap(3);
            if (berr <= tol) break;
            if (it >= max_sweeps || !(rmax <= stall * prev) || rmax != rmax) {
                fb = true;
                break;
            }
            prev = rmax;
            for (int i = tid; i < n; i += LX_WG_THREADS) vf[i] = (float)(slot.r[i] / rmax);
            __syncthreads();
            lx_b_sweep<float>(F, ld, Wf, slot.ldw, n, 0, vf, smem);
            lx_b_sweep<float>(F, ld, Wf, slot.ldw, n, 1, vf, smem);
            for (int i = tid; i < n; i += LX_WG_THREADS) S.X[i] += rmax * (double)vf[i];
            __syncthreads();
            out.sweeps = it + 1;
            clk.lap(4);
        }
    }
    if (fb) {
        // FP64 factor of the same system (the FP32 factor's storage is reused)
        double *D = reinterpret_cast<double *>(slot.F);
        double *v = slot.v;
        lx_b_lower_to<double>(S.A, S.lda, n, D, ld);
        int info64 = 0;
        lx_unit_potrf<double>(D, ld, n, slot.W, slot.ldw, &info64, smem);
        info = __builtin_amdgcn_readfirstlane(info64);
        for (int i = tid; i < n; i += LX_WG_THREADS) v[i] = S.B[i];
        __syncthreads();
        lx_b_sweep<double>(D, ld, slot.W, slot.ldw, n, 0, v, smem);
        lx_b_sweep<double>(D, ld, slot.W, slot.ldw, n, 1, v, smem);
        for (int i = tid; i < n; i += LX_WG_THREADS) S.X[i] = v[i];
        __syncthreads();
        lx_b_resid(S.A, S.lda, n, S.X, S.B, slot.r, true, smem, rmax, xmax, anorm);
        berr = rmax / (anorm * xmax);
        clk.lap(5);
    }
    out.fallback = fb ? 1 : 0;
    out.info = info;
    out.berr = berr;
    if (tid == 0) *stat = out;
    __syncthreads();
}

__device__ __forceinline__ lx_bslot lx_b_slot(char *ws, int64_t slot_bytes, int nmax)
{
    const int64_t nm = ((int64_t)nmax + 15) & ~(int64_t)15;
    char *base = ws + (int64_t)blockIdx.x * slot_bytes;
    lx_bslot s;
    s.F = base;
    s.ldw = nm + 64;
    s.W = reinterpret_cast<double *>(base + nm * nm * 8);
    s.r = s.W + s.ldw * 64;
    s.v = s.r + nm;
    return s;
}

// Bytes of one workspace slot for systems of order <= nmax.
__host__ __device__ inline int64_t lx_b_slot_bytes(int nmax)
{
    const int64_t nm = ((int64_t)nmax + 15) & ~(int64_t)15;
    return nm * nm * 8 + (nm + 64) * 64 * 8 + 2 * nm * 8;
}

// Persistent launch: workgroups take systems order[0], order[1], ... from
// *ticket (zero at launch) until all count are done.  MINB = resident
// workgroups per CU the registers are budgeted for (1: no spills, 2: twice
// the latency hiding); the autotuner picks.  Bit c of pw_mask: systems of
// order 64c+1 .. 64c+64 take 128-column factor panels (else 64).
template <int MINB>
__global__ void __launch_bounds__(LX_WG_THREADS, MINB)
lx_batch_ir_kernel(const lx_bsys *__restrict__ sys, const int *__restrict__ order, int count, char *ws,
                   int64_t slot_bytes, int nmax, unsigned *ticket, lx_bstat *stat, int max_sweeps, double tol,
                   double stall, long long *prof, unsigned pw_mask)
{
    __shared__ __attribute__((aligned(16))) char smem[LX_BATCH_SMEM];
    __shared__ int s_t;
    __shared__ long long s_prof[LX_B_PROF_WORDS];
    const lx_bslot slot = lx_b_slot(ws, slot_bytes, nmax);
    lx_b_clock clk;
    clk.acc = prof ? s_prof : nullptr;
    if (prof && lx_tid() < LX_B_PROF_WORDS) s_prof[lx_tid()] = 0;  // ordered by the loop's first barrier
    for (;;) {
        if (lx_tid() == 0) s_t = (int)atomicAdd(ticket, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_t);
        __syncthreads();
        if (t >= count) break;
        const int k = order[t];
        const lx_bsys S = sys[k];
        const int pw = (S.n > 0 && ((pw_mask >> ((S.n - 1) / 64)) & 1u)) ? 128 : 64;
        lx_b_solve_one(S, slot, stat + k, max_sweeps, tol, stall, smem, clk, pw);
        if (prof && lx_tid() == 0) s_prof[6] += 1;
    }
    if (prof && lx_tid() == 0)
        for (int q = 0; q < LX_B_PROF_WORDS; ++q) prof[(int64_t)blockIdx.x * LX_B_PROF_WORDS + q] = s_prof[q];
}

__global__ void lx_batch_reset_kernel(unsigned *ticket) { ticket[0] = 0u; }

#ifndef LX_BATCH_KERNELS_ONLY
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#include "lx_internal.hpp"

static_assert(sizeof(lx_bstat) == sizeof(linompss_batch_status), "status layout");

struct linompss_batch_plan {
    int count = 0, nmax = 0, cus = 0;
    int wg_per_cu = 1, order = 0, grid = 0, use_graph = 1;
    unsigned pw_mask = 0;  // bit c: orders 64c+1 .. 64c+64 factor in 128-column panels
    int max_sweeps = 30;
    double tol = 1e-14, stall = 0.5, tune_ms = 0.0;
    int64_t slot_bytes = 0;
    lx_bsys *d_sys = nullptr;
    int *d_order[2] = {nullptr, nullptr};  // [0] largest first, [1] as given
    lx_bstat *d_stat = nullptr;
    unsigned *d_ticket = nullptr;
    char *d_ws = nullptr;
    long long *d_prof = nullptr;  // phase profile (LINOMPSS_BATCH_PROFILE set at plan creation)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Autotuned (workgroups per CU, order, panel mask) by batch signature:
// device, count, largest order and the total factor work.
struct lx_bt_choice {
    int wg, order;
    unsigned pw_mask;
    bool pw_tuned;  // pw_mask was measured (not a fixed width of the plan that stored it)
};
std::mutex lx_bt_mu;
std::map<std::tuple<int, int, int, long long>, lx_bt_choice> lx_bt_cache;

// The panel rule without measurements: 128-column panels above order 256.
unsigned lx_b_default_mask() { return 0xF0u; }

// One persistent launch over cnt systems taken from d_ord.
int lx_b_launch_ex(linompss_batch_plan *p, int wg, const int *d_ord, int cnt, unsigned mask, hipStream_t s)
{
    const int grid = std::min(cnt, std::max(1, wg * p->cus));
    hipLaunchKernelGGL(lx_batch_reset_kernel, dim3(1), dim3(1), 0, s, p->d_ticket);
    if (wg == 1)
        hipLaunchKernelGGL(lx_batch_ir_kernel<1>, dim3(grid), dim3(LX_WG_THREADS), 0, s, p->d_sys, d_ord, cnt,
                           p->d_ws, p->slot_bytes, p->nmax, p->d_ticket, p->d_stat, p->max_sweeps, p->tol, p->stall,
                           p->d_prof, mask);
    else
        hipLaunchKernelGGL(lx_batch_ir_kernel<2>, dim3(grid), dim3(LX_WG_THREADS), 0, s, p->d_sys, d_ord, cnt,
                           p->d_ws, p->slot_bytes, p->nmax, p->d_ticket, p->d_stat, p->max_sweeps, p->tol, p->stall,
                           p->d_prof, mask);
    LX_HIP_TRY(hipGetLastError());
    return 0;
}

int lx_b_launch(linompss_batch_plan *p, int wg, int ord, hipStream_t s)
{
    return lx_b_launch_ex(p, wg, p->d_order[ord], p->count, p->pw_mask, s);
}

void lx_b_release(linompss_batch_plan *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->d_sys);
    (void)hipFree(p->d_order[0]);
    (void)hipFree(p->d_order[1]);
    (void)hipFree(p->d_stat);
    (void)hipFree(p->d_ticket);
    (void)hipFree(p->d_ws);
    (void)hipFree(p->d_prof);
    delete p;
}

// The panel width of each size class, measured on the batch's own systems:
// for every class with at least 16 systems, the class alone (largest first)
// is run with 64- and with 128-column panels under the chosen workgroups per
// CU, and a width at least 2 % faster sets the class's bit (the runs write
// X).  Classes with fewer systems, or closer timings, keep the default rule.
int lx_b_tune_panels(linompss_batch_plan *p, const int *n, int wg, unsigned *mask_out, const int *ord_desc)
{
    hipStream_t s = nullptr;
    LX_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int *d_sub = nullptr;
    int rc = 0;
    unsigned mask = lx_b_default_mask();
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMalloc(&d_sub, sizeof(int) * std::max(1, p->count)) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    for (int c = 0; c < LX_BATCH_NMAX / 64 && rc == 0; ++c) {
        std::vector<int> sub;
        for (int i = 0; i < p->count; ++i) {
            const int k = ord_desc[i];
            if (n[k] > 64 * c && n[k] <= 64 * c + 64) sub.push_back(k);
        }
        if (sub.size() < 16) continue;
        if (hipMemcpy(d_sub, sub.data(), sizeof(int) * sub.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = LINOMPSS_ERR_HIP;
            break;
        }
        float t[2] = {0.f, 0.f};
        for (int v = 0; v < 2 && rc == 0; ++v) {
            const unsigned m = v ? (mask | (1u << c)) : (mask & ~(1u << c));
            if ((rc = lx_b_launch_ex(p, wg, d_sub, (int)sub.size(), m, s)) != 0) break;  // warm-up
            if (hipEventRecord(e0, s) != hipSuccess || (rc = lx_b_launch_ex(p, wg, d_sub, (int)sub.size(), m, s)) != 0 ||
                hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t[v], e0, e1) != hipSuccess) {
                if (rc == 0) rc = LINOMPSS_ERR_HIP;
                break;
            }
            p->tune_ms += t[v];
        }
        // a width wins by 2 % or more; closer than that the default rule stays
        if (rc == 0 && t[1] < 0.98f * t[0]) mask |= 1u << c;
        if (rc == 0 && t[0] < 0.98f * t[1]) mask &= ~(1u << c);
    }
    if (d_sub) (void)hipFree(d_sub);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc == 0) *mask_out = mask;
    return rc;
}

// Times each (workgroups per CU, order) candidate on the batch itself (the
// runs write X) and returns the fastest.
int lx_b_tune(linompss_batch_plan *p, bool tune_wg, bool tune_order, int *wg_out, int *ord_out)
{
    hipStream_t s = nullptr;
    LX_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    float best = INFINITY;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = LINOMPSS_ERR_HIP;
    const int wgs[2] = {1, 2}, ords[2] = {0, 1};
    for (int a = 0; a < 2 && rc == 0; ++a) {
        const int wg = tune_wg ? wgs[a] : p->wg_per_cu;
        if (!tune_wg && a > 0) break;
        for (int o = 0; o < 2 && rc == 0; ++o) {
            const int ord = tune_order ? ords[o] : p->order;
            if (!tune_order && o > 0) break;
            float ms = 0.f;
            if ((rc = lx_b_launch(p, wg, ord, s)) != 0) break;  // warm-up
            if (hipEventRecord(e0, s) != hipSuccess || (rc = lx_b_launch(p, wg, ord, s)) != 0 ||
                hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                if (rc == 0) rc = LINOMPSS_ERR_HIP;
                break;
            }
            p->tune_ms += ms;
            if (ms < best) {
                best = ms;
                *wg_out = wg;
                *ord_out = ord;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return rc;
}

}  // namespace

extern "C" int linompss_batch_plan_create(int count, const int *n, const double *const *dA, const int64_t *lda,
                                          const double *const *dB, double *const *dX,
                                          const linompss_batch_opts *opts, linompss_batch_plan **plan)
{
    if (!plan) return -8;
    *plan = nullptr;
    if (count < 0) return -1;
    linompss_batch_opts o = {30, 1e-14, 0.5, 0, -1, 1, 0};
    if (opts) o = *opts;
    if (o.wg_per_cu < 0 || o.wg_per_cu > 2 || o.order < -1 || o.order > 1 || o.max_sweeps < 0) return -7;
    if (o.panel != 0 && o.panel != 64 && o.panel != 128) return -7;
    if (count > 0 && (!n || !dA || !lda || !dB || !dX)) return -2;
    int nmax = 0;
    double work = 0.0;
    std::vector<lx_bsys> hs(count);
    for (int i = 0; i < count; ++i) {
        if (n[i] < 0 || n[i] > LX_BATCH_NMAX) return -2;
        if (n[i] > 0 && !dA[i]) return -3;
        if (lda[i] < std::max(1, n[i])) return -4;
        if (n[i] > 0 && !dB[i]) return -5;
        if (n[i] > 0 && !dX[i]) return -6;
        hs[i] = lx_bsys{dA[i], dB[i], dX[i], lda[i], n[i], 0};
        nmax = std::max(nmax, n[i]);
        work += (double)n[i] * n[i] * n[i];
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LINOMPSS_ERR_NODEV;
    linompss_batch_plan *p = new linompss_batch_plan();
    p->count = count;
    p->nmax = std::max(1, nmax);
    p->max_sweeps = o.max_sweeps;
    p->tol = o.tol;
    p->stall = o.stall_ratio;
    p->use_graph = o.use_graph ? 1 : 0;
    p->wg_per_cu = o.wg_per_cu ? o.wg_per_cu : 1;
    p->order = o.order < 0 ? 0 : o.order;
    p->pw_mask = o.panel == 64 ? 0u : o.panel == 128 ? 0xFFu : lx_b_default_mask();
    p->slot_bytes = lx_b_slot_bytes(p->nmax);
    auto fail = [&](int rc) {
        lx_b_release(p);
        return rc;
    };
    if (hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || p->cus < 1)
        return fail(LINOMPSS_ERR_HIP);
    if (count == 0) {
        *plan = p;
        return 0;
    }
    // LPT order: largest system first (stable)
    std::vector<int> ord_desc(count), ord_given(count);
    std::iota(ord_given.begin(), ord_given.end(), 0);
    ord_desc = ord_given;
    std::stable_sort(ord_desc.begin(), ord_desc.end(), [&](int a, int b) { return n[a] > n[b]; });
    const int gmax = std::min(count, 2 * p->cus);
    if (hipMalloc(&p->d_sys, sizeof(lx_bsys) * count) != hipSuccess ||
        hipMalloc(&p->d_order[0], sizeof(int) * count) != hipSuccess ||
        hipMalloc(&p->d_order[1], sizeof(int) * count) != hipSuccess ||
        hipMalloc(&p->d_stat, sizeof(lx_bstat) * count) != hipSuccess ||
        hipMalloc(&p->d_ticket, 256) != hipSuccess ||
        hipMalloc(&p->d_ws, (size_t)p->slot_bytes * gmax) != hipSuccess)
        return fail(LINOMPSS_ERR_NOMEM);
    if (getenv("LINOMPSS_BATCH_PROFILE") &&
        hipMalloc(&p->d_prof, sizeof(long long) * LX_B_PROF_WORDS * gmax) != hipSuccess)
        return fail(LINOMPSS_ERR_NOMEM);
    if (hipMemcpy(p->d_sys, hs.data(), sizeof(lx_bsys) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_order[0], ord_desc.data(), sizeof(int) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_order[1], ord_given.data(), sizeof(int) * count, hipMemcpyHostToDevice) != hipSuccess)
        return fail(LINOMPSS_ERR_HIP);
    // schedule: given, cached for this batch signature, or measured now (the
    // panel mask too, when it is left per size class)
    const bool tune_wg = o.wg_per_cu == 0, tune_order = o.order < 0;
    if (tune_wg || tune_order) {
        const bool tune_pw = o.panel == 0;
        // total work in buckets of about 1 % (log scale)
        const long long sig = std::llround(std::log2(work + 1.0) * 64.0);
        const auto key = std::make_tuple(dev, count, nmax, sig);
        bool hit = false, pw_hit = false;
        {
            std::lock_guard<std::mutex> g(lx_bt_mu);
            auto it = lx_bt_cache.find(key);
            if (it != lx_bt_cache.end()) {
                if (tune_wg) p->wg_per_cu = it->second.wg;
                if (tune_order) p->order = it->second.order;
                if (tune_pw && it->second.pw_tuned) {
                    p->pw_mask = it->second.pw_mask;
                    pw_hit = true;
                }
                hit = true;
            }
        }
        if (!hit || (tune_pw && !pw_hit)) {
            int wg = p->wg_per_cu, od = p->order, rc = 0;
            if (!hit) {
                if ((rc = lx_b_tune(p, tune_wg, tune_order, &wg, &od)) != 0) return fail(rc);
                p->wg_per_cu = wg;
                p->order = od;
            }
            bool pw_tuned = false;
            if (tune_pw) {
                unsigned m = p->pw_mask;
                if ((rc = lx_b_tune_panels(p, n, p->wg_per_cu, &m, ord_desc.data())) != 0) return fail(rc);
                p->pw_mask = m;
                pw_tuned = true;
            }
            std::lock_guard<std::mutex> g(lx_bt_mu);
            auto it = lx_bt_cache.find(key);
            if (it == lx_bt_cache.end())
                lx_bt_cache[key] = {p->wg_per_cu, p->order, p->pw_mask, pw_tuned};
            else if (pw_tuned) {  // keep the stored schedule, add the measured panels
                it->second.pw_mask = p->pw_mask;
                it->second.pw_tuned = true;
            }
        }
    }
    p->grid = std::min(count, std::max(1, p->wg_per_cu * p->cus));
    if (p->use_graph) {
        // the launch pair (ticket reset, persistent kernel) as one graph
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return fail(LINOMPSS_ERR_HIP);
        int rc = 0;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = LINOMPSS_ERR_HIP;
        if (rc == 0) {
            rc = lx_b_launch(p, p->wg_per_cu, p->order, cs);
            if (hipStreamEndCapture(cs, &p->graph) != hipSuccess && rc == 0) rc = LINOMPSS_ERR_HIP;
        }
        if (rc == 0 && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = LINOMPSS_ERR_HIP;
        (void)hipStreamDestroy(cs);
        if (rc) return fail(rc);
    }
    *plan = p;
    return 0;
}

extern "C" int linompss_batch_run(linompss_batch_plan *p, void *stream, linompss_batch_status *st)
{
    if (!p) return -1;
    if (p->count == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (p->use_graph) {
        LX_HIP_TRY(hipGraphLaunch(p->exec, s));
    } else {
        const int rc = lx_b_launch(p, p->wg_per_cu, p->order, s);
        if (rc) return rc;
    }
    if (st) LX_HIP_TRY(hipMemcpyAsync(st, p->d_stat, sizeof(lx_bstat) * p->count, hipMemcpyDeviceToHost, s));
    LX_HIP_TRY(hipStreamSynchronize(s));
    if (p->d_prof) {
        std::vector<long long> h((size_t)LX_B_PROF_WORDS * p->grid);
        LX_HIP_TRY(hipMemcpy(h.data(), p->d_prof, sizeof(long long) * h.size(), hipMemcpyDeviceToHost));
        double tot[LX_B_PROF_WORDS] = {}, all = 0.0;
        for (int g = 0; g < p->grid; ++g)
            for (int q = 0; q < LX_B_PROF_WORDS; ++q) tot[q] += (double)h[(size_t)g * LX_B_PROF_WORDS + q];
        for (int q = 0; q < 6; ++q) all += tot[q];
        static const char *nm[6] = {"copy", "factor", "solve0", "resid", "corr", "fallback"};
        fprintf(stderr, "[linompss batch profile] grid=%d systems=%.0f busy per WG %.3f ms:", p->grid, tot[6],
                all / 1e5 / p->grid);
        for (int q = 0; q < 6; ++q) fprintf(stderr, " %s %.1f%%", nm[q], 100.0 * tot[q] / (all > 0 ? all : 1.0));
        fprintf(stderr, "\n");
    }
    return 0;
}

extern "C" int linompss_batch_plan_panels(const linompss_batch_plan *p, unsigned *mask)
{
    if (!p) return -1;
    if (!mask) return -2;
    *mask = p->pw_mask;
    return 0;
}

extern "C" int linompss_batch_plan_info(const linompss_batch_plan *p, int *grid, int *wg_per_cu, int *order,
                                        double *tune_ms)
{
    if (!p) return -1;
    if (grid) *grid = p->grid;
    if (wg_per_cu) *wg_per_cu = p->wg_per_cu;
    if (order) *order = p->order;
    if (tune_ms) *tune_ms = p->tune_ms;
    return 0;
}

extern "C" int linompss_batch_plan_destroy(linompss_batch_plan *p)
{
    if (p) {
        (void)hipDeviceSynchronize();
        lx_b_release(p);
    }
    return 0;
}

extern "C" int linompss_dposv_ir_batched_dev(int count, const int *n, const double *const *dA, const int64_t *lda,
                                             const double *const *dB, double *const *dX, void *stream,
                                             const linompss_batch_opts *opts, linompss_batch_status *st)
{
    linompss_batch_opts o = {30, 1e-14, 0.5, 1, 0, 0, 0};  // one launch, no tuning, no graph
    if (opts) o = *opts;
    linompss_batch_plan *p = nullptr;
    int rc = linompss_batch_plan_create(count, n, dA, lda, dB, dX, &o, &p);
    if (rc) return rc;
    rc = linompss_batch_run(p, stream, st);
    linompss_batch_plan_destroy(p);
    return rc;
}
#endif  // LX_BATCH_KERNELS_ONLY
