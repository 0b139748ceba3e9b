// lx_sweep.hip -- blocked triangular sweeps on the device: L Y = B and
// L^T X = Y for a lower factor L, nrhs <= 16 right-hand sides.
//
// Same arithmetic as the reference's bblas_dltrsm lower branches
// (src/blocked_blas/bblas_trsm.h:119-160): for each row block, subtract the
// products with the blocks already solved (task_dgemm), then solve with the
// diagonal block (task_dtrsm).  The reference runs these as a chain of OmpSs
// tasks over b-row blocks; on the GPU the sweep is bandwidth- and
// latency-bound (L is read once, n^2/2 elements, and each block depends on
// the previous one), so it runs as one launch with 64-row blocks:
//   * a workgroup takes the next block by ticket (blocks are handed out in
//     dependency order, so every wait is on a block already held by a running
//     workgroup);
//   * its 4 wavefronts stream the L blocks of its row (forward) or column
//     (backward), each waiting on the ready flag of the block it consumes;
//     the backward sweep transposes each 64 x 16 slab through LDS so that the
//     global reads stay coalesced;
//   * wavefront 0 solves the 64 x 64 diagonal block (lane = row, shuffles
//     broadcast the solved entries), stores the block and publishes its flag.
// All branches that steer a wavefront are wave-uniform.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "linompss.h"
#include "lx_device.hpp"
#include "lx_internal.hpp"

#define LX_SW_NR 16

__device__ __forceinline__ int lx_sw_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T>
__device__ __forceinline__ void lx_sw_wait(const int *flag)
{
    while (lx_sw_uniform(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0)
        __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// trans = 0: forward L Y = B;  trans = 1: backward L^T X = Y.  X is
// overwritten in place.  flags[nblk] and *ticket must be zero on entry.
//
// The 64 x 64 diagonal block is read into registers (16 values per thread,
// coalesced) before the workgroup waits on anything, and moved to LDS for the
// solve, so the serial part of the chain touches no global memory.
template <typename T, int TRANS>
__global__ void __launch_bounds__(256) lx_sweep_kernel(int n, int nrhs, const T *__restrict__ L, int64_t ldl,
                                                       T *__restrict__ X, int64_t ldx, int *flags, int *ticket)
{
    constexpr int KC = 128 / sizeof(T);  // backward: L rows per LDS slab (one 128-byte line per column)
    constexpr int TILE = 4 * 64 * (KC + 1), DIAG = 64 * 65;
    // per wave: the solved block it consumes, then (after its last block) its partial sums
    __shared__ T sx[4][LX_SW_NR][64];
    // backward transpose slabs during the accumulation, the diagonal block for the solve
    __shared__ T sbuf[(TRANS ? TILE : 0) > DIAG ? (TRANS ? TILE : 0) : DIAG];
    __shared__ int s_blk;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nblk = (n + 63) / 64;
    if (w == 0) {
        const int t = lx_sw_uniform(atomicAdd(ticket, lane == 0 ? 1 : 0));
        s_blk = t;  // every lane of wave 0 stores the same value
    }
    __syncthreads();
    const int t = lx_sw_uniform(s_blk);
    const int ib = TRANS ? nblk - 1 - t : t;
    const int r0 = ib * 64;
    const int hb = min(64, n - r0);
    const int row = r0 + min(lane, hb - 1);  // clamped: lanes past hb recompute the last row, never store

    // diagonal block -> registers: thread (lane, w) holds column 4*q + w, row lane
    T dreg[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int col = min(4 * q + w, hb - 1);
        dreg[q] = L[(int64_t)(r0 + col) * ldl + row];
    }

    T acc[LX_SW_NR];
#pragma unroll
    for (int c = 0; c < LX_SW_NR; ++c) acc[c] = T(0);

    // contributions of the blocks solved before this one, oldest first (forward:
    // j = 0, 1, ...; backward: j = nblk-1, nblk-2, ...) so that each wave only
    // ever waits for the newest block at the end of its list
    const int nj = TRANS ? nblk - 1 - ib : ib;
    for (int q = w; q < nj; q += 4) {
        const int j = TRANS ? nblk - 1 - q : q;
        lx_sw_wait<T>(&flags[j]);
        const int c0 = j * 64, hj = min(64, n - c0);
        // stage x_j (solved rows c0 .. c0+hj) in LDS
        for (int c = 0; c < nrhs; ++c) sx[w][c][lane] = lane < hj ? X[(int64_t)c * ldx + c0 + lane] : T(0);
        __builtin_amdgcn_wave_barrier();
        if (!TRANS) {
            // acc[r] += sum_k L[row, c0+k] x_j[k]  (column k of the block: 64 contiguous rows)
            const T *Lb = L + (int64_t)c0 * ldl + row;
            if (hj == 64) {
                for (int k0 = 0; k0 < 64; k0 += 8) {
                    T l[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) l[u] = Lb[(int64_t)(k0 + u) * ldl];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
#pragma unroll
                        for (int c = 0; c < LX_SW_NR; ++c)
                            if (c < nrhs) acc[c] += l[u] * sx[w][c][k0 + u];
                }
            } else {
                for (int k = 0; k < hj; ++k) {
                    const T l = Lb[(int64_t)k * ldl];
#pragma unroll
                    for (int c = 0; c < LX_SW_NR; ++c)
                        if (c < nrhs) acc[c] += l * sx[w][c][k];
                }
            }
        } else {
            // acc[r] += sum_k L[c0+k, r0+r] x_j[k]: block (j, ib) read along its columns
            // (one 128-byte line per column and instruction), transposed through LDS
            T(*stile)[KC + 1] = reinterpret_cast<T(*)[KC + 1]>(sbuf + w * 64 * (KC + 1));
            constexpr int CPI = 64 / KC;  // columns per load instruction
            for (int k0 = 0; k0 < hj; k0 += KC) {
                const int kk = lane % KC, cq = lane / KC;
                T v[KC];  // KC >= 16: 64 / CPI loads in flight
#pragma unroll
                for (int q = 0; q < 64 / CPI; ++q) {
                    const int cc = q * CPI + cq;
                    const int colg = r0 + min(cc, hb - 1), rowg = c0 + min(k0 + kk, hj - 1);
                    v[q] = L[(int64_t)colg * ldl + rowg];
                }
#pragma unroll
                for (int q = 0; q < 64 / CPI; ++q) stile[q * CPI + cq][kk] = v[q];
                __builtin_amdgcn_wave_barrier();
                const int kend = min(KC, hj - k0);
                for (int k = 0; k < kend; ++k) {
                    const T l = stile[lane][k];
#pragma unroll
                    for (int c = 0; c < LX_SW_NR; ++c)
                        if (c < nrhs) acc[c] += l * sx[w][c][k0 + k];
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // sx[w] is this wave's own buffer: safe to reuse for its partial sums
#pragma unroll
    for (int c = 0; c < LX_SW_NR; ++c) sx[w][c][lane] = acc[c];
    __syncthreads();  // all transpose slabs consumed: sbuf now takes the diagonal block
    T(*sdiag)[65] = reinterpret_cast<T(*)[65]>(sbuf);  // sdiag[r][c] = L[r0+r, r0+c]
#pragma unroll
    for (int q = 0; q < 16; ++q) sdiag[lane][4 * q + w] = dreg[q];
    __syncthreads();
    if (w == 0) {
        // right-hand side of this block minus the contributions
        T s[LX_SW_NR];
#pragma unroll
        for (int c = 0; c < LX_SW_NR; ++c)
            s[c] = c < nrhs ? X[(int64_t)c * ldx + row] -
                                  ((sx[0][c][lane] + sx[1][c][lane]) + (sx[2][c][lane] + sx[3][c][lane]))
                            : T(0);
        const int lr = min(lane, hb - 1);
        // diagonal block solve, lane = row
        if (!TRANS) {
            for (int jj = 0; jj < hb; ++jj) {
                const T piv = sdiag[jj][jj];
                const T ljj = sdiag[lr][jj];  // L[row, r0+jj]
#pragma unroll
                for (int c = 0; c < LX_SW_NR; ++c) {
                    if (c < nrhs) {
                        const T xj = __shfl(s[c], jj) / piv;
                        s[c] = lane == jj ? xj : (lane > jj ? s[c] - ljj * xj : s[c]);
                    }
                }
            }
        } else {
            for (int jj = hb - 1; jj >= 0; --jj) {
                const T piv = sdiag[jj][jj];
                const T ljj = sdiag[jj][lr];  // L[r0+jj, row] = (L^T)[row, r0+jj]
#pragma unroll
                for (int c = 0; c < LX_SW_NR; ++c) {
                    if (c < nrhs) {
                        const T xj = __shfl(s[c], jj) / piv;
                        s[c] = lane == jj ? xj : (lane < jj ? s[c] - ljj * xj : s[c]);
                    }
                }
            }
        }
        for (int c = 0; c < nrhs; ++c)
            if (lane < hb) X[(int64_t)c * ldx + r0 + lane] = s[c];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[ib], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

#ifndef LX_SWEEP_KERNELS_ONLY
template <typename T>
static int sweeps(int n, int nrhs, const T *L, int64_t ldl, T *X, int64_t ldx, int which, hipStream_t s)
{
    const int nblk = (n + 63) / 64;
    int *ctl = (int *)lx_scratch(3, sizeof(int) * (size_t)(2 * nblk + 64));
    if (!ctl) return LINOMPSS_ERR_NOMEM;
    for (int c0 = 0; c0 < nrhs; c0 += LX_SW_NR) {
        const int nc = std::min(LX_SW_NR, nrhs - c0);
        if (which & 1) {
            lx_fill_i32(ctl, 0, nblk + 64, s);
            hipLaunchKernelGGL((lx_sweep_kernel<T, 0>), dim3(nblk), dim3(256), 0, s, n, nc, L, ldl, X + c0 * ldx, ldx,
                               ctl + 64, ctl);
        }
        if (which & 2) {
            lx_fill_i32(ctl, 0, nblk + 64, s);
            hipLaunchKernelGGL((lx_sweep_kernel<T, 1>), dim3(nblk), dim3(256), 0, s, n, nc, L, ldl, X + c0 * ldx, ldx,
                               ctl + 64, ctl);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

// which: 1 forward, 2 backward, 3 both.  Enqueued on the stream (no sync).
int lx_sweeps_dev(int dtype, int n, int nrhs, const void *L, int64_t ldl, void *X, int64_t ldx, int which,
                  hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return 0;
    if (dtype == 0) return sweeps<double>(n, nrhs, (const double *)L, ldl, (double *)X, ldx, which, s);
    return sweeps<float>(n, nrhs, (const float *)L, ldl, (float *)X, ldx, which, s);
}
#endif  // LX_SWEEP_KERNELS_ONLY
