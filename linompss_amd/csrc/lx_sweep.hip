// lx_sweep.hip -- blocked triangular sweeps on the device: L Y = B and
// L^T X = Y for a lower factor L, nrhs <= 16 right-hand sides per launch.
//
// Same arithmetic as the reference's bblas_dltrsm lower branches
// (src/blocked_blas/bblas_trsm.h:119-160): for each row block, subtract the
// products with the blocks already solved (task_dgemm), then solve with the
// diagonal block (task_dtrsm).  The reference runs these as a chain of OmpSs
// tasks over b-row blocks.  On the GPU a sweep reads L once (n^2/2 elements,
// HBM-bound) along a chain of dependent diagonal solves (latency-bound), so
// one persistent launch splits it into 64-row steps:
//   * the products with blocks solved two or more steps earlier are partial
//     sums computed by many workgroups in parallel (C blocks per unit), each
//     written to its own slot and added in a fixed order (deterministic);
//   * the solve unit of a step computes the inverse of its 64 x 64 diagonal
//     block and loads the block coupling it to the previous step before it
//     waits, so the serial part of the chain is one 64 x 64 product with the
//     previous step's x, the sum of the partials and one 64 x 64 product with
//     the inverse.
// All branches that steer a wavefront are wave-uniform; every wait is bounded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <vector>

#include "linompss.h"
#include "lx_device.hpp"
#include "lx_internal.hpp"

#define LX_SW_NR 16

// waves per SIMD a kernel's registers are budgeted for (the CPU build of the
// kernels defines it away)
#ifndef LX_WAVES_PER_EU
#define LX_WAVES_PER_EU(w) __attribute__((amdgpu_waves_per_eu(w, w)))
#endif

__device__ __forceinline__ int lx_sw_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Work units of one sweep over nb 64-row blocks, in "step" order s = 0..nb-1
// (block s forward, block nb-1-s backward):
//   S(s)     the solve unit of step s: b_s minus the partial sums of step s,
//            minus the products with the blocks solved at steps s-GAP..s-1
//            (each subtracted as soon as that step is published), times the
//            inverse of the diagonal block; publishes x_s through the mailbox;
//   P(s, c)  c >= 1: the partial sum of step s over the steps
//            [max(0, s-GAP-cC), s-GAP-(c-1)C); publishes by counting pcnt[s].
// A partial sum never reads the GAP newest steps, so its units run GAP-1
// steps ahead of the chain of solve units, and the counter's release and
// acquire stay off the chain.  Units are listed by the newest step they read
// (S(s): s-1, P(s, c): s-GAP-1-(c-1)C), S first within a key.  Every wait is
// on a unit listed earlier, so a persistent grid that takes units in list
// order cannot deadlock.  Unit code: (s << 16) | c.
#define LX_SW_GAP 3  // the solve unit's direct products (s-3, s-2) use sD and sA
__host__ __device__ constexpr int lx_sw_nparts(int s, int C, int gap = LX_SW_GAP)
{
    return s > gap ? (s - gap + C - 1) / C : 0;
}
static inline void lx_sw_build_units(int nb, int C, std::vector<int> &units, std::vector<int> &rowbase,
                                     int gap = LX_SW_GAP)
{
    units.clear();
    rowbase.assign(nb + 1, 0);
    for (int s = 0; s < nb; ++s) rowbase[s + 1] = rowbase[s] + lx_sw_nparts(s, C, gap);
    for (int key = -1; key <= nb - 2; ++key) {
        units.push_back((key + 1) << 16);
        if (key < 0) continue;
        for (int c = 1;; ++c) {
            const long long st = (long long)key + gap + 1 + (long long)(c - 1) * C;
            if (st >= nb) break;
            units.push_back((int)(st << 16) | c);
        }
    }
}

// Poll until *p >= need (wave-uniform; every lane of the calling wave must be
// active).  Gives up after spin_ticks of the constant wall clock, then sets
// *abort so that every other wait ends too.  Followed by an acquire fence.
__device__ __forceinline__ void lx_sw_wait_ge(const int *p, int need, int *abort, long long spin_ticks)
{
    if (lx_sw_uniform(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
        const long long t0 = wall_clock64();
        while (lx_sw_uniform(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
            if (lx_sw_uniform(__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) break;
            if (lx_sw_uniform(wall_clock64() - t0 > spin_ticks ? 1 : 0)) {  // one decision per wave
                atomicOr(abort, (threadIdx.x & 63) == 0 ? 1 : 0);  // read-modify-write: seen at once
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// *p += 1 once the workgroup's stores are visible device-wide: each wave waits
// for its own stores, then wave 0 alone issues the agent-scope release
// read-modify-write (one L2 write-back for the workgroup, not one per wave).
__device__ __forceinline__ void lx_sw_publish(int *p)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (lx_wave() == 0) {  // wave-uniform; lane 0 adds 1, the others 0
        // the release as a fence plus an explicit wait, then a relaxed add: with a
        // release read-modify-write the compiler dropped the wait between the L2
        // write-back and the atomic as redundant (nothing else was outstanding),
        // letting the counter overtake the data
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(p, (threadIdx.x & 63) == 0 ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Mailbox of solved values: one slot per element of x, all-ones bits (a
// negative quiet NaN with a full payload) until the solve unit writes it with
// an atomic exchange, which is performed at the device's shared level at once.
// A consumer polls its element until the bits change: the value arrives with
// its own "ready" signal, one memory round trip and no release fence on the
// chain of solve units.  Arithmetic keeps an input NaN's payload, so a NaN in
// B or L can reach the mailbox with exactly those bits: the writer replaces
// that one pattern by the canonical quiet NaN, and the sweep propagates NaN
// instead of leaving its readers waiting for a value that looks unwritten.
template <typename T> struct lx_sw_bits;
template <> struct lx_sw_bits<float> {
    typedef unsigned int U;
    static constexpr U qnan = 0x7fc00000u;
};
template <> struct lx_sw_bits<double> {
    typedef unsigned long long U;
    static constexpr U qnan = 0x7ff8000000000000ull;
};

template <typename T>
__device__ __forceinline__ void lx_sw_mb_put(T *mb, T v)
{
    typedef typename lx_sw_bits<T>::U U;
    U bits = __builtin_bit_cast(U, v);
    if (bits == ~U(0)) bits = lx_sw_bits<T>::qnan;
    __hip_atomic_exchange((U *)mb, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-lane wait (a lane loop with no barrier inside); bounded like every wait.
template <typename T>
__device__ __forceinline__ T lx_sw_mb_get(const T *mb, int *abort, long long spin_ticks)
{
    typedef typename lx_sw_bits<T>::U U;
    U v = __hip_atomic_load((const U *)mb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == ~U(0)) {
        const long long t0 = wall_clock64();
        for (;;) {
            v = __hip_atomic_load((const U *)mb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v != ~U(0)) break;
            if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (wall_clock64() - t0 > spin_ticks) {
                atomicOr(abort, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return __builtin_bit_cast(T, v);
}

// One load of a mailbox slot (its bits; all-ones = not yet written), so that
// a lane can issue the loads of several slots before it polls any of them.
template <typename T>
__device__ __forceinline__ typename lx_sw_bits<T>::U lx_sw_mb_peek(const T *mb)
{
    return __hip_atomic_load((const typename lx_sw_bits<T>::U *)mb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The value of a peeked slot, polling (bounded) while it read empty.
template <typename T>
__device__ __forceinline__ T lx_sw_mb_finish(const T *mb, typename lx_sw_bits<T>::U v, int *abort,
                                             long long spin_ticks)
{
    typedef typename lx_sw_bits<T>::U U;
    return v == ~U(0) ? lx_sw_mb_get<T>(mb, abort, spin_ticks) : __builtin_bit_cast(T, v);
}

// trans = 0: forward L Y = B;  trans = 1: backward L^T X = Y.  X (n x nrhs,
// nrhs <= NR <= 16) is overwritten in place.  ctl = [ticket, abort, pad...,
// unused[nb], pcnt[nb]] must be zero on entry; part holds rowbase[nb] * 16 * 64
// values; mb (leading dimension ldmb >= 64 nb, nrhs columns) must hold
// all-ones bits on entry.
template <typename T, int TRANS, int NR>
__global__ void __launch_bounds__(256) lx_sweep_kernel(int n, int nrhs, const T *__restrict__ L, int64_t ldl,
                                                       T *__restrict__ X, int64_t ldx, const int *__restrict__ units,
                                                       int nunits, const int *__restrict__ rowbase, int C,
                                                       T *__restrict__ part, int *ctl, T *__restrict__ mb,
                                                       int64_t ldmb, long long spin_ticks)
{
    constexpr int KC = 128 / sizeof(T);  // backward partial sums: L rows per LDS slab
    constexpr int S_BYTES = 4 * 64 * 65 * (int)sizeof(T);
    constexpr int P_BYTES = (4 * 64 * (KC + 1) + 4 * NR * 64) * (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) char smem[S_BYTES > P_BYTES ? S_BYTES : P_BYTES];
    __shared__ T svec[3][64][NR + 1];  // S unit: [0] = b, then y; [1], [2] = solved blocks' x
    __shared__ int s_unit;
    const int nb = (n + 63) / 64;
    int *ticket = ctl, *abort = ctl + 1, *pcnt = ctl + 64 + nb;
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave();
    for (;;) {
        if (w == 0) {
            const int t = lx_sw_uniform(atomicAdd(ticket, lane == 0 ? 1 : 0));
            s_unit = t < nunits ? units[t] : -1;  // every lane stores the same value
        }
        __syncthreads();
        const int code = lx_sw_uniform(s_unit);
        __syncthreads();
        if (code < 0) break;
        const int s = code >> 16, c = code & 0xffff;
        const int ib = TRANS ? nb - 1 - s : s;
        const int r0 = ib * 64, hb = min(64, n - r0);
        if (c == 0) {
            // ---------------- S(s) ----------------
            // sA: the diagonal block, then the block coupling step s-2;
            // sZ: its inverse; sC: the block coupling step s-1; sD: step s-3
            T(*sA)[65] = reinterpret_cast<T(*)[65]>(smem);
            T(*sZ)[65] = reinterpret_cast<T(*)[65]>(smem + 64 * 65 * sizeof(T));
            T(*sC)[65] = reinterpret_cast<T(*)[65]>(smem + 2 * 64 * 65 * sizeof(T));
            T(*sD)[65] = reinterpret_cast<T(*)[65]>(smem + 3 * 64 * 65 * sizeof(T));
            // coupling block with step s-d into sB: sB[r][k] = L[r0 + r][ip*64 + k]
            // (forward) or L[ip*64 + k][r0 + r] (backward); reads run down columns
            auto load_coupling = [&](T(*sB)[65], int d) {
                const int ip = TRANS ? ib + d : ib - d;
                const int hp = min(64, n - ip * 64);
                for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
                    const int i0 = e % 64, i1 = e / 64;
                    if (!TRANS)
                        sB[i0][i1] = i0 < hb ? L[(int64_t)(ip * 64 + i1) * ldl + r0 + i0] : T(0);
                    else
                        sB[i1][i0] = i0 < hp ? L[(int64_t)(r0 + i1) * ldl + ip * 64 + i0] : T(0);
                }
            };
            // nothing before the waits depends on another unit
            for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
                const int r = e % 64, k = e / 64;
                sA[r][k] = (r < hb && k < hb) ? L[(int64_t)(r0 + k) * ldl + r0 + r] : (r == k ? T(1) : T(0));
                sZ[r][k] = T(0);
            }
            for (int e = tid; e < 64 * NR; e += LX_WG_THREADS) {
                const int r = e % 64, cc = e / 64;
                svec[0][r][cc] = (r < hb && cc < nrhs) ? X[(int64_t)cc * ldx + r0 + r] : T(0);
            }
            if (s >= 1) load_coupling(sC, 1);
            if (s >= 3) load_coupling(sD, 3);
            __syncthreads();
            lx_tri_inv64<T>(sA, sZ);  // ends with a barrier
            // M1 = op(Z) C1 over sC (op = transpose for the backward sweep), so that
            // once step s-1 is published one product remains: x = op(Z) y - M1 x_{s-1};
            // thread (row lane, columns 16w..16w+15)
            if (s >= 1) {
                T m[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) m[q] = T(0);
                for (int j = 0; j < 64; ++j) {
                    const T z = TRANS ? sZ[j][lane] : sZ[lane][j];
#pragma unroll
                    for (int q = 0; q < 16; ++q) m[q] += z * sC[j][16 * w + q];
                }
                __syncthreads();  // every wave has read C1
#pragma unroll
                for (int q = 0; q < 16; ++q) sC[lane][16 * w + q] = m[q];
            }
            if (s >= 2) load_coupling(sA, 2);
            // b minus the partial sums (in part order; they only read steps <=
            // s-GAP-1), lane = row, wave w owning rhs w, w+4, ...; the host picks C
            // so that a step has at most 8 parts, and all loads of a column are
            // issued together.  Wave 0 waits for the parts' counter and subtracts
            // its columns at once, while waves 1..3 take x_{s-3} and x_{s-2} from the
            // mailbox (svec[1], svec[2]): the counter's wait and the two mailbox
            // reads overlap.
            const int need = lx_sw_nparts(s, C);
            const T *pp = part + (size_t)rowbase[s] * LX_SW_NR * 64;
            auto sub_parts = [&](int cc) {
                T y = svec[0][lane][cc];
                for (int q0 = 0; q0 < need; q0 += 8) {
                    T v[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        v[q] = q0 + q < need ? pp[((size_t)(q0 + q) * LX_SW_NR + cc) * 64 + lane] : T(0);
#pragma unroll
                    for (int q = 0; q < 8; ++q) y -= v[q];
                }
                svec[0][lane][cc] = y;
            };
            if (w == 0) {  // wave-uniform
                if (need > 0) {
                    lx_sw_wait_ge(&pcnt[s], need, abort, spin_ticks);
                    for (int cc = 0; cc < nrhs; cc += 4) sub_parts(cc);
                }
            } else {
                for (int e = tid - 64; e < 2 * 64 * NR; e += LX_WG_THREADS - 64) {
                    const int h = e / (64 * NR), e2 = e % (64 * NR), k = e2 % 64, cc = e2 / 64;
                    const int d = 3 - h;
                    if (s < d) continue;
                    const int ip = TRANS ? ib + d : ib - d;
                    const int hp = min(64, n - ip * 64);
                    svec[1 + h][k][cc] = (k < hp && cc < nrhs)
                                             ? lx_sw_mb_get<T>(mb + (int64_t)cc * ldmb + ip * 64 + k, abort, spin_ticks)
                                             : T(0);
                }
            }
            __syncthreads();  // wave 0 past the counter's acquire; svec[1], svec[2] filled
            if (need > 0 && w > 0)  // wave-uniform
                for (int cc = w; cc < nrhs; cc += 4) sub_parts(cc);
            // the products with the blocks solved at steps s-3 and s-2
            for (int d = LX_SW_GAP; d >= 2; --d) {
                if (s < d) continue;  // uniform
                T(*sB)[65] = d == 3 ? sD : sA;
                const int h = 3 - d;
                for (int cc = w; cc < nrhs; cc += 4) {  // wave-uniform
                    T a0 = T(0), a1 = T(0);
#pragma unroll 8
                    for (int k = 0; k < 64; k += 2) {
                        a0 += sB[lane][k] * svec[1 + h][k][cc];
                        a1 += sB[lane][k + 1] * svec[1 + h][k + 1][cc];
                    }
                    svec[0][lane][cc] -= a0 + a1;
                }
            }
            // u = op(Z) y for this wave's right-hand sides, kept in registers
            constexpr int NU = (NR + 3) / 4;
            T u[NU];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < NU; ++t) {
                const int cc = w + 4 * t;
                T x0 = T(0), x1 = T(0);
                if (cc < nrhs) {  // wave-uniform; column cc is this wave's own
#pragma unroll 8
                    for (int k = 0; k < 64; k += 2) {
                        x0 += (TRANS ? sZ[k][lane] : sZ[lane][k]) * svec[0][k][cc];
                        x1 += (TRANS ? sZ[k + 1][lane] : sZ[lane][k + 1]) * svec[0][k + 1][cc];
                    }
                }
                u[t] = x0 + x1;
            }
            if (s >= 1) {  // uniform
                // the chain: x = u - M1 x_{s-1}.  Wave w takes entries 16w..16w+15 of
                // x_{s-1} from the mailbox and forms the product over those 16 columns
                // of M1 for every row; the four partial sums meet in LDS (over sA and
                // sZ, free once every wave has formed u)
                const int ip = TRANS ? ib + 1 : ib - 1;
                const int hp = min(64, n - ip * 64);
                T(*spart)[64][NR + 1] = reinterpret_cast<T(*)[64][NR + 1]>(smem);
                __syncthreads();  // u formed in every wave, svec[1] free, M1 in sC
                for (int e = lane; e < 16 * NR; e += 64) {
                    const int k = 16 * w + (e & 15), cc = e >> 4;
                    svec[1][k][cc] = (k < hp && cc < nrhs)
                                         ? lx_sw_mb_get<T>(mb + (int64_t)cc * ldmb + ip * 64 + k, abort, spin_ticks)
                                         : T(0);
                }
                __builtin_amdgcn_wave_barrier();
                for (int cc = 0; cc < nrhs; ++cc) {  // uniform
                    T a0 = T(0), a1 = T(0);
#pragma unroll
                    for (int q = 0; q < 16; q += 2) {
                        a0 += sC[lane][16 * w + q] * svec[1][16 * w + q][cc];
                        a1 += sC[lane][16 * w + q + 1] * svec[1][16 * w + q + 1][cc];
                    }
                    spart[w][lane][cc] = a0 + a1;
                }
                __syncthreads();
#pragma unroll
                for (int t = 0; t < NU; ++t) {
                    const int cc = w + 4 * t;
                    if (cc < nrhs) {  // wave-uniform
                        u[t] -= (spart[0][lane][cc] + spart[1][lane][cc]) + (spart[2][lane][cc] + spart[3][lane][cc]);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < NU; ++t) {
                const int cc = w + 4 * t;
                if (cc < nrhs && lane < hb) {
                    X[(int64_t)cc * ldx + r0 + lane] = u[t];
                    lx_sw_mb_put<T>(mb + (int64_t)cc * ldmb + r0 + lane, u[t]);
                }
            }
        } else {
            // ---------------- P(s, c) ----------------
            T(*stile)[KC + 1] = reinterpret_cast<T(*)[KC + 1]>(smem + w * 64 * (KC + 1) * sizeof(T));
            T(*sx)[64] = reinterpret_cast<T(*)[64]>(smem + (4 * 64 * (KC + 1) + w * NR * 64) * sizeof(T));
            const int t_hi = s - LX_SW_GAP - (c - 1) * C, t_lo = max(0, t_hi - C);
            const int row = r0 + min(lane, hb - 1);  // clamped: lanes past hb never store
            T acc[NR];
#pragma unroll
            for (int cc = 0; cc < NR; ++cc) acc[cc] = T(0);
            // several right-hand sides: the block products on the matrix cores
            // (four 16-row tiles of 16 columns, 16 MFMAs of K = 4 each per block);
            // lane l supplies op(L)[16 tt + l%16][4 kk + l/16] (forward: from
            // global memory; backward: from the slab transposed in LDS) and
            // x[4 kk + l/16][l%16]
            constexpr bool MM = NR > 1;
            typename lx_mfma<T>::acc macc[4];
#pragma unroll
            for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                for (int i = 0; i < 4; ++i) macc[tt][i] = T(0);
            for (int t = t_lo + w; t < t_hi; t += 4) {  // wave-uniform
                const int jb = TRANS ? nb - 1 - t : t;
                const int c0 = jb * 64, hj = min(64, n - c0);
                if constexpr (MM && !TRANS) {
                    // the block of L before the wait (block jb < ib is full; rows past hb clamped)
                    const int lr = lane & 15, lk = lane >> 4;
                    T a[4][16];
#pragma unroll
                    for (int tt = 0; tt < 4; ++tt) {
                        const T *Lb = L + (int64_t)(c0 + lk) * ldl + r0 + min(16 * tt + lr, hb - 1);
#pragma unroll
                        for (int kk = 0; kk < 16; ++kk) a[tt][kk] = Lb[(int64_t)(4 * kk) * ldl];
                    }
                    for (int cc = 0; cc < nrhs; ++cc)
                        sx[cc][lane] = lane < hj ? lx_sw_mb_get<T>(mb + (int64_t)cc * ldmb + c0 + lane, abort, spin_ticks)
                                                 : T(0);
                    __builtin_amdgcn_wave_barrier();
#pragma unroll
                    for (int kk = 0; kk < 16; ++kk) {
                        const T bv = lr < nrhs ? sx[lr][4 * kk + lk] : T(0);
#pragma unroll
                        for (int tt = 0; tt < 4; ++tt) macc[tt] = lx_mfma<T>::mac(a[tt][kk], bv, macc[tt]);
                    }
                    __builtin_amdgcn_wave_barrier();
                    continue;
                }
                // the block of L is read before the wait: only x_j depends on step t
                T lv[64];
                if (!TRANS) {
                    // lv[k] = L[row, c0 + k]: column k of the block is 64 contiguous rows
                    // (forward: block jb < ib is full)
                    const T *Lb = L + (int64_t)c0 * ldl + row;
#pragma unroll
                    for (int k = 0; k < 64; ++k) lv[k] = Lb[(int64_t)k * ldl];
                } else {
                    // slab k0 = KC*m: lv[m*(64/CPI) + q] = L[c0 + k0 + kk, r0 + q*CPI + cq], one
                    // 128-byte line per column and instruction
                    constexpr int CPI = 64 / KC;
                    const int kk = lane % KC, cq = lane / KC;
#pragma unroll
                    for (int m = 0; m < 64 / KC; ++m)
#pragma unroll
                        for (int q = 0; q < 64 / CPI; ++q) {
                            const int colg = r0 + min(q * CPI + cq, hb - 1), rowg = c0 + min(m * KC + kk, hj - 1);
                            lv[m * (64 / CPI) + q] = L[(int64_t)colg * ldl + rowg];
                        }
                }
                for (int cc = 0; cc < nrhs; ++cc)
                    sx[cc][lane] = lane < hj ? lx_sw_mb_get<T>(mb + (int64_t)cc * ldmb + c0 + lane, abort, spin_ticks)
                                             : T(0);
                __builtin_amdgcn_wave_barrier();
                if (!TRANS) {
#pragma unroll
                    for (int k = 0; k < 64; ++k)
#pragma unroll
                        for (int cc = 0; cc < NR; ++cc)
                            if (cc < nrhs) acc[cc] += lv[k] * sx[cc][k];
                } else {
                    // transpose each slab through LDS: stile[col][kk]
                    constexpr int CPI = 64 / KC;
                    const int kk = lane % KC, cq = lane / KC;
#pragma unroll
                    for (int m = 0; m < 64 / KC; ++m) {
#pragma unroll
                        for (int q = 0; q < 64 / CPI; ++q) stile[q * CPI + cq][kk] = lv[m * (64 / CPI) + q];
                        __builtin_amdgcn_wave_barrier();
                        // rows past hj: sx is zero there
                        if constexpr (MM) {
                            const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
                            for (int k4 = 0; k4 < KC; k4 += 4) {
                                const T bv = lr < nrhs ? sx[lr][m * KC + k4 + lk] : T(0);
#pragma unroll
                                for (int tt = 0; tt < 4; ++tt)
                                    macc[tt] = lx_mfma<T>::mac(stile[16 * tt + lr][k4 + lk], bv, macc[tt]);
                            }
                        } else {
#pragma unroll 8
                            for (int k = 0; k < KC; ++k) {
                                const T l = stile[lane][k];
#pragma unroll
                                for (int cc = 0; cc < NR; ++cc)
                                    if (cc < nrhs) acc[cc] += l * sx[cc][m * KC + k];
                            }
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            // reduce the four waves in a fixed order through sx (each wave's own buffer)
            if constexpr (MM) {
                const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
                for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (lr < NR) sx[lr][16 * tt + lx_mfma<T>::row(lk, i)] = macc[tt][i];
            } else {
#pragma unroll
                for (int cc = 0; cc < NR; ++cc) sx[cc][lane] = acc[cc];
            }
            __syncthreads();
            T(*sx0)[NR][64] = reinterpret_cast<T(*)[NR][64]>(smem + 4 * 64 * (KC + 1) * sizeof(T));
            T *pp = part + ((size_t)rowbase[s] + (c - 1)) * LX_SW_NR * 64;
            for (int cc = w; cc < nrhs; cc += 4) {
                const T v = (sx0[0][cc][lane] + sx0[1][cc][lane]) + (sx0[2][cc][lane] + sx0[3][cc][lane]);
                pp[(size_t)cc * 64 + lane] = lane < hb ? v : T(0);
            }
            lx_sw_publish(&pcnt[s]);
        }
    }
}

// ---------------------------------------------------------------------------
// One right-hand side: 128-row steps over precomputed step matrices.
//
// The chain of the sweep above costs about one cross-XCD mailbox round trip
// per 64-row step.  With one right-hand side the chain can take 128-row steps
// when the solve unit's matrices come from a pre-pass over L alone: for each
// 128-row block i, with D = L_ii (padded with the identity past n),
//   forward:  Z = D^-1,   M1 = D^-1 L_{i,i-1},     M2 = D^-1 L_{i,i-2}
//   backward: Z = D^-T,   M1 = D^-T L_{i+1,i}^T,   M2 = D^-T L_{i+2,i}^T
// so that x_i = Z (b_i - partial sums) - M2 x_{i-/+2} - M1 x_{i-/+1}: the
// serial part of a step is one 128 x 128 product with the previous step's x,
// from registers.  The pre-pass reads only L, so a refinement loop runs it
// once per factor and reuses it for every sweep.  Layout of the pre-pass
// output: for direction d (0 forward, 1 backward) and block i, three
// column-major 128 x 128 matrices Z, M1, M2 at ((d nb + i) 3 + m) 128^2
// (a coupling past either end of L is stored as zeros).
#define LX_S2 128
#define LX_S2_GAP 2  // partial sums skip the two newest steps (the solve unit's M2 and M1)
#define LX_S2_MAT (LX_S2 * LX_S2)

// Partial sums of step s, of sizes growing geometrically from C0 steps: part
// c = 1, 2, ... covers the steps [hi - C0 2^(c-1), hi) with
// hi = s - gap - C0 (2^(c-1) - 1), clipped at 0.  The part the solve unit
// waits for last (c = 1, the newest steps) is short, so it is done soon after
// its newest step is; the larger parts end at older steps and are done long
// before they are needed.
__host__ __device__ inline void lx_s2_part(int s, int c, int C0, int &lo, int &hi)
{
    const int sz = C0 << (c - 1);
    hi = s - LX_S2_GAP - (sz - C0);
    lo = hi - sz > 0 ? hi - sz : 0;
}
__host__ __device__ inline int lx_s2_nparts(int s, int C0)
{
    int m = 0;
    for (int done = 0; done < s - LX_S2_GAP; ++m) done += C0 << m;
    return m;
}
// Unit list of a 128-row sweep over nb blocks: S(s) keyed by s - 1, P(s, c)
// by the newest step it reads (hi - 1), S first within a key, then by s.
// Every wait is on a unit listed earlier, so a persistent grid that takes
// units in list order cannot deadlock.  rowbase[s]: index of step s's first
// part.
static inline void lx_s2_build_units(int nb, int C0, std::vector<int> &units, std::vector<int> &rowbase)
{
    rowbase.assign(nb + 1, 0);
    std::vector<std::array<int, 3>> v;  // key, 0 = S / 1 = P, code
    for (int s = 0; s < nb; ++s) {
        const int np = lx_s2_nparts(s, C0);
        rowbase[s + 1] = rowbase[s] + np;
        v.push_back({s - 1, 0, s << 16});
        for (int c = 1; c <= np; ++c) {
            int lo, hi;
            lx_s2_part(s, c, C0, lo, hi);
            v.push_back({hi - 1, 1, (s << 16) | c});
        }
    }
    std::stable_sort(v.begin(), v.end(), [](const std::array<int, 3> &a, const std::array<int, 3> &b) {
        return a[0] != b[0] ? a[0] < b[0] : a[1] < b[1];
    });
    units.clear();
    for (const auto &e : v) units.push_back(e[2]);
}

// One workgroup per block: grid nb, one launch per direction (DIR 0 forward,
// 1 backward).  Every load loop has a constant trip count with all its loads
// issued before the LDS stores, so a thread has its 16 or 64 loads in flight
// at once.
template <typename T, int DIR>
__global__ void __launch_bounds__(256) lx_sw2_prep_kernel(int n, const T *__restrict__ L, int64_t ldl,
                                                          T *__restrict__ prep)
{
    constexpr int WS = LX_S2 + 4;                                  // row stride of sW: 16-byte aligned rows
    __shared__ T sZ[LX_S2][LX_S2 + 1];                             // D^-1 (lower)
    __shared__ __attribute__((aligned(16))) T sW[LX_S2][WS];       // scratch, then a coupling block
    const int nb = (n + LX_S2 - 1) / LX_S2, ib = blockIdx.x;
    const int r0 = ib * LX_S2, hb = min(LX_S2, n - r0);
    const int tid = lx_tid();
    T(*sA)[65] = reinterpret_cast<T(*)[65]>(&sW[0][0]);
    T(*sB)[65] = reinterpret_cast<T(*)[65]>(&sW[0][0] + 64 * 65);
    // a 64 x 64 block of L at (row o1, column o2) of the diagonal block, rows
    // past hb (and, on the diagonal, columns past hb) as the identity's
    auto load64 = [&](T(*dst)[65], int o1, int o2, bool diag) {
        T v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = tid + i * LX_WG_THREADS, r = e % 64, k = e / 64;
            v[i] = (o1 + r < hb && o2 + k < hb) ? L[(int64_t)(r0 + o2 + k) * ldl + r0 + o1 + r]
                                                : (diag && r == k ? T(1) : T(0));
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = tid + i * LX_WG_THREADS;
            dst[e % 64][e / 64] = v[i];
        }
    };
    for (int e = tid; e < LX_S2 * (LX_S2 + 1); e += LX_WG_THREADS) (&sZ[0][0])[e] = T(0);
    // the two 64 x 64 diagonal blocks of D, inverted in turn
    for (int h = 0; h < 2; ++h) {
        load64(sA, 64 * h, 64 * h, true);
        for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) sB[e % 64][e / 64] = T(0);
        __syncthreads();
        lx_tri_inv64<T>(sA, sB);  // ends with a barrier
        for (int e = tid; e < 64 * 64; e += LX_WG_THREADS) {
            const int r = e % 64, k = e / 64;
            sZ[64 * h + r][64 * h + k] = sB[r][k];
        }
        __syncthreads();
    }
    // Z21 = -Z22 (L21 Z11): sB = L21 (rows past n zero), sA = L21 Z11
    load64(sB, 64, 0, false);
    __syncthreads();
    {
        const int i = tid & 63, j0 = 16 * (tid >> 6);
        T acc[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = T(0);
        for (int k = 0; k < 64; ++k) {
            const T a = sB[i][k];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] += a * sZ[k][j0 + q];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) sA[i][j0 + q] = acc[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = T(0);
        for (int k = 0; k < 64; ++k) {
            const T a = sZ[64 + i][64 + k];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] += a * sA[k][j0 + q];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) sZ[64 + i][j0 + q] = -acc[q];
    }
    __syncthreads();
    T *pz = prep + (size_t)(DIR * nb + ib) * 3 * LX_S2_MAT;
    for (int e = tid; e < LX_S2_MAT; e += LX_WG_THREADS) {
        const int r = e % LX_S2, c = e / LX_S2;
        pz[e] = DIR ? sZ[c][r] : sZ[r][c];
    }
    // M_d = op(Z) C_d: C[k][c] = L[r0 + k, ip*128 + c] (forward, ip = ib - d) or
    // L[ip*128 + c, r0 + k] (backward, ip = ib + d); thread: row r, 64 columns.
    // Load e = tid + 256 i of the block: lo = tid % 128 (k forward, c
    // backward), hi = tid / 128 + 2 i, consecutive lanes along a column of L.
#pragma unroll 1
    for (int d = 1; d <= 2; ++d) {
        const int ip = DIR ? ib + d : ib - d;
        T *pm = pz + (size_t)d * LX_S2_MAT;
        if (ip < 0 || ip >= nb) {  // uniform
            for (int e = tid; e < LX_S2_MAT; e += LX_WG_THREADS) pm[e] = T(0);
            continue;
        }
        const int hp = min(LX_S2, n - ip * LX_S2);
        const int lo = tid & (LX_S2 - 1), hi0 = tid >> 7;
        const bool lo_ok = DIR ? lo < hp : lo < hb;
        const int hi_lim = DIR ? hb : hp;
        const T *src = DIR ? L + (int64_t)(r0 + hi0) * ldl + ip * LX_S2 + lo : L + (int64_t)(ip * LX_S2 + hi0) * ldl + r0 + lo;
        const int64_t step = 2 * ldl;
        T v[LX_S2_MAT / LX_WG_THREADS];
#pragma unroll
        for (int i = 0; i < LX_S2_MAT / LX_WG_THREADS; ++i)
            v[i] = (lo_ok && hi0 + 2 * i < hi_lim) ? src[i * step] : T(0);
        __syncthreads();  // sW free
#pragma unroll
        for (int i = 0; i < LX_S2_MAT / LX_WG_THREADS; ++i) {
            if (!DIR) sW[lo][hi0 + 2 * i] = v[i];
            else sW[hi0 + 2 * i][lo] = v[i];
        }
        __syncthreads();
        const int r = tid & (LX_S2 - 1), c0 = 64 * (tid >> 7);
        T acc[64];
#pragma unroll
        for (int q = 0; q < 64; ++q) acc[q] = T(0);
        for (int k = 0; k < LX_S2; ++k) {
            const T z = DIR ? sZ[k][r] : sZ[r][k];
            const T *wk = &sW[k][c0];
#pragma unroll
            for (int q = 0; q < 64; ++q) acc[q] += z * wk[q];
        }
#pragma unroll
        for (int q = 0; q < 64; ++q) pm[(size_t)(c0 + q) * LX_S2 + r] = acc[q];
    }
}

// Padded LDS index of a 128-vector read by two half-waves at once (entries k
// and 64 + k land in different banks).
__device__ __forceinline__ int lx_s2_ix(int k) { return k + 4 * (k >> 6); }

// trans = 0: forward L y = x;  trans = 1: backward L^T x = y; one column X,
// overwritten in place.  Units of lx_s2_build_units (128-row steps, a gap of
// 2, parts of C, 2C, 4C, ... steps).  Partial sums are mailboxes too: P(s, c) writes its 128 values
// with atomic exchanges and S(s) polls them by value, so the path from the
// newest step a part reads to the solve unit that adds it is two memory
// round trips (no counter, no release fence), within the two steps of slack
// the gap gives.  ctl = [ticket, abort, ...] zero on entry; part
// (rowbase[nb] * 128) and mb (nb * 128) all-ones bits on entry; prep from
// lx_sw2_prep_kernel over the same L.
template <typename T, int TRANS, int WPE>
__global__ void __launch_bounds__(256) LX_WAVES_PER_EU(WPE) lx_sweep2_kernel(int n, const T *__restrict__ L, int64_t ldl, T *__restrict__ X,
                                                        const T *__restrict__ prep, const int *__restrict__ units,
                                                        int nunits, const int *__restrict__ rowbase, int C,
                                                        T *__restrict__ part, int *ctl, T *__restrict__ mb,
                                                        long long spin_ticks)
{
    constexpr int KC = 128 / sizeof(T);  // backward partial sums: L rows per LDS slab
    constexpr int CPI = 64 / KC;
    constexpr int V = LX_S2 + 8;  // padded 128-vector (lx_s2_ix)
    __shared__ __attribute__((aligned(16))) T stile[4][64][KC + 1];
    __shared__ __attribute__((aligned(16))) T sxw[4][V];  // per wave: x of a block (P: 64, S: x_{s-1})
    __shared__ __attribute__((aligned(16))) T sy[V], sx2[V];
    __shared__ T sred[2][LX_S2];
    __shared__ int s_unit;
    const int nb = (n + LX_S2 - 1) / LX_S2;
    int *ticket = ctl, *abort = ctl + 1;
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave();
    for (;;) {
        if (w == 0) {
            const int t = lx_sw_uniform(atomicAdd(ticket, lane == 0 ? 1 : 0));
            s_unit = t < nunits ? units[t] : -1;  // every lane stores the same value
        }
        __syncthreads();
        const int code = lx_sw_uniform(s_unit);
        __syncthreads();
        if (code < 0) break;
        const int s = code >> 16, c = code & 0xffff;
        const int ib = TRANS ? nb - 1 - s : s;
        const int r0 = ib * LX_S2, hb = min(LX_S2, n - r0);
        if (c == 0) {
            // ---------------- S(s) ----------------
            // thread: row rl of the block, columns 64 hc .. 64 hc + 63 of Z, M1, M2
            const int rl = 32 * w + (lane & 31), hc = lane >> 5;
            const T *pz = prep + (size_t)(TRANS * nb + ib) * 3 * LX_S2_MAT + (size_t)(64 * hc) * LX_S2 + rl;
            T mz[64], m1[64], m2[64];
#pragma unroll
            for (int q = 0; q < 64; ++q) {
                mz[q] = pz[q * LX_S2];
                m2[q] = pz[2 * LX_S2_MAT + q * LX_S2];
                m1[q] = pz[LX_S2_MAT + q * LX_S2];
            }
            // waves 0, 1: y = b minus the partial sums (each slot's load issued
            // before any is polled, added in part order); waves 2, 3: x_{s-2}
            const int need = lx_s2_nparts(s, C);
            if (tid < LX_S2) {  // waves 0 and 1
                T y = tid < hb ? X[r0 + tid] : T(0);
                const T *pp = part + (size_t)rowbase[s] * LX_S2 + tid;
                for (int q0 = 0; q0 < need; q0 += 8) {
                    typename lx_sw_bits<T>::U v[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (q0 + q < need) v[q] = lx_sw_mb_peek<T>(pp + (size_t)(q0 + q) * LX_S2);
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (q0 + q < need) y -= lx_sw_mb_finish<T>(pp + (size_t)(q0 + q) * LX_S2, v[q], abort, spin_ticks);
                }
                sy[lx_s2_ix(tid)] = y;
            } else {
                const int k = 64 * (w - 2) + lane;
                T v = T(0);
                if (s >= 2) {
                    const int ip = TRANS ? ib + 2 : ib - 2;
                    if (k < min(LX_S2, n - ip * LX_S2)) v = lx_sw_mb_get<T>(mb + ip * LX_S2 + k, abort, spin_ticks);
                }
                sx2[lx_s2_ix(k)] = v;
            }
            __syncthreads();
            // u = Z y - M2 x_{s-2}; the two half-waves' sums meet by a lane swap
            const T *yv = sy + 68 * hc, *x2v = sx2 + 68 * hc;  // = lx_s2_ix(64 hc + q) - q
            T a0 = T(0), a1 = T(0), b0 = T(0), b1 = T(0);
#pragma unroll
            for (int q = 0; q < 64; q += 2) {
                a0 += mz[q] * yv[q];
                a1 += mz[q + 1] * yv[q + 1];
                b0 += m2[q] * x2v[q];
                b1 += m2[q + 1] * x2v[q + 1];
            }
            T a = a0 + a1, bb = b0 + b1;
            a += __shfl_xor(a, 32);
            bb += __shfl_xor(bb, 32);
            T u = a - bb;
            if (s >= 1) {  // uniform: the chain, x = u - M1 x_{s-1}
                const int ip = TRANS ? ib + 1 : ib - 1;
                const int hp = min(LX_S2, n - ip * LX_S2);
                T *sx1 = sxw[w];
                const T *p0 = mb + ip * LX_S2 + lane, *p1 = p0 + 64;
                typename lx_sw_bits<T>::U e0 = 0, e1 = 0;  // zero bits: the value 0 past hp
                if (lane < hp) e0 = lx_sw_mb_peek<T>(p0);
                if (64 + lane < hp) e1 = lx_sw_mb_peek<T>(p1);
                sx1[lx_s2_ix(lane)] = lx_sw_mb_finish<T>(p0, e0, abort, spin_ticks);
                sx1[lx_s2_ix(64 + lane)] = lx_sw_mb_finish<T>(p1, e1, abort, spin_ticks);
                __builtin_amdgcn_wave_barrier();
                const T *x1v = sx1 + 68 * hc;
                T c0 = T(0), c1 = T(0);
#pragma unroll
                for (int q = 0; q < 64; q += 2) {
                    c0 += m1[q] * x1v[q];
                    c1 += m1[q + 1] * x1v[q + 1];
                }
                T cs = c0 + c1;
                cs += __shfl_xor(cs, 32);
                u -= cs;
            }
            if (hc == 0 && rl < hb) {
                X[r0 + rl] = u;
                lx_sw_mb_put<T>(mb + r0 + rl, u);
            }
        } else {
            // ---------------- P(s, c) ----------------
            // steps [t_lo, t_hi); wave w: rows 64 (w >> 1) .. of the block against
            // columns 64 (w & 1) .. of each earlier block
            int t_lo, t_hi;
            lx_s2_part(s, c, C, t_lo, t_hi);
            const int rh = w >> 1, ch = w & 1;
            const int hr = min(64, hb - 64 * rh);
            T acc = T(0);
            T *sx = sxw[w];
            if (hr > 0) {  // wave-uniform
                const int row = r0 + 64 * rh + min(lane, hr - 1);  // clamped: rows past hr never stored
                for (int t = t_lo; t < t_hi; ++t) {
                    const int jb = TRANS ? nb - 1 - t : t;
                    const int c0 = jb * LX_S2 + 64 * ch, hj = min(64, n - c0);
                    if (hj <= 0) continue;  // uniform (a ragged last block, backward)
                    // the block of L before the wait: only x depends on step t
                    T lv[64];
                    if (!TRANS) {
                        const T *Lb = L + (int64_t)c0 * ldl + row;
#pragma unroll
                        for (int k = 0; k < 64; ++k) lv[k] = Lb[(int64_t)k * ldl];
                    } else {
                        const int kk = lane % KC, cq = lane / KC;
#pragma unroll
                        for (int m = 0; m < 64 / KC; ++m)
#pragma unroll
                            for (int q = 0; q < 64 / CPI; ++q) {
                                const int colg = r0 + 64 * rh + min(q * CPI + cq, hr - 1);
                                const int rowg = c0 + min(m * KC + kk, hj - 1);
                                lv[m * (64 / CPI) + q] = L[(int64_t)colg * ldl + rowg];
                            }
                    }
                    sx[lane] = lane < hj ? lx_sw_mb_get<T>(mb + c0 + lane, abort, spin_ticks) : T(0);
                    __builtin_amdgcn_wave_barrier();
                    if (!TRANS) {
                        T p0 = T(0), p1 = T(0);
#pragma unroll
                        for (int k = 0; k < 64; k += 2) {
                            p0 += lv[k] * sx[k];
                            p1 += lv[k + 1] * sx[k + 1];
                        }
                        acc += p0 + p1;
                    } else {
                        const int kk = lane % KC, cq = lane / KC;
#pragma unroll
                        for (int m = 0; m < 64 / KC; ++m) {
#pragma unroll
                            for (int q = 0; q < 64 / CPI; ++q) stile[w][q * CPI + cq][kk] = lv[m * (64 / CPI) + q];
                            __builtin_amdgcn_wave_barrier();
                            T p0 = T(0), p1 = T(0);
#pragma unroll
                            for (int k = 0; k < KC; k += 2) {  // rows past hj: x is zero there
                                p0 += stile[w][lane][k] * sx[m * KC + k];
                                p1 += stile[w][lane][k + 1] * sx[m * KC + k + 1];
                            }
                            acc += p0 + p1;
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            sred[ch][64 * rh + lane] = acc;
            __syncthreads();
            T *pp = part + ((size_t)rowbase[s] + (c - 1)) * LX_S2;
            if (tid < LX_S2) lx_sw_mb_put<T>(pp + tid, tid < hb ? sred[0][tid] + sred[1][tid] : T(0));
        }
    }
}

// Up to 16 right-hand sides (nrhs columns of X, ld ldx; the mailbox and the
// parts hold 16 columns, ld ldmb = nb * 128 and 128): the same units and step
// matrices as lx_sweep2_kernel, the products on the matrix cores
// (v_mfma_f32_16x16x4f32, lane l: A[l%16][l/16], B[l/16][l%16]).  The solve
// unit holds Z, M1 and M2 as A operands in registers (wave w: rows 32w ..
// 32w + 31 as two 16-row tiles, 32 k-steps each); Y = b - parts, x_{s-2} and
// x_{s-1} pass through LDS as B operands ([column][row], row stride 132: the
// 64 lanes of a read hit 64 banks).  A partial-sum unit's wave (rows 64 rh ..,
// columns 64 ch .. of each earlier block) runs four 16-row tiles over its
// block's 64 columns, L from global memory (forward) or from the slab
// transposed in LDS (backward).
template <typename T, int TRANS>
__global__ void __launch_bounds__(256) LX_WAVES_PER_EU(1) lx_sweep2m_kernel(
    int n, int nrhs, const T *__restrict__ L, int64_t ldl, T *__restrict__ X, int64_t ldx, const T *__restrict__ prep,
    const int *__restrict__ units, int nunits, const int *__restrict__ rowbase, int C, T *__restrict__ part, int *ctl,
    T *__restrict__ mb, int64_t ldmb, long long spin_ticks)
{
    typedef typename lx_sw_bits<T>::U U;
    typedef lx_mfma<T> MF;
    constexpr int KC = 128 / sizeof(T);
    constexpr int CPI = 64 / KC;
    constexpr int VS = LX_S2 + 4;  // row stride of the [column][row] B operands
    __shared__ __attribute__((aligned(16))) T stile[4][64][KC + 1];
    __shared__ __attribute__((aligned(16))) T sxp[4][16][68];  // P: per wave, x of a 64-row block
    __shared__ __attribute__((aligned(16))) T sY[16][VS], sX2[16][VS], sX1[16][VS];
    __shared__ T sred[2][16][LX_S2];
    __shared__ int s_unit;
    const int nb = (n + LX_S2 - 1) / LX_S2;
    int *ticket = ctl, *abort = ctl + 1;
    const int tid = lx_tid(), lane = tid & 63, w = lx_wave(), lr = lane & 15, lk = lane >> 4;
    for (;;) {
        if (w == 0) {
            const int t = lx_sw_uniform(atomicAdd(ticket, lane == 0 ? 1 : 0));
            s_unit = t < nunits ? units[t] : -1;  // every lane stores the same value
        }
        __syncthreads();
        const int code = lx_sw_uniform(s_unit);
        __syncthreads();
        if (code < 0) break;
        const int s = code >> 16, c = code & 0xffff;
        const int ib = TRANS ? nb - 1 - s : s;
        const int r0 = ib * LX_S2, hb = min(LX_S2, n - r0);
        if (c == 0) {
            // ---------------- S(s) ----------------
            const T *pz = prep + (size_t)(TRANS * nb + ib) * 3 * LX_S2_MAT + 32 * w + lr + (size_t)lk * LX_S2;
            T mz[2][32], m1[2][32], m2[2][32];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int kk = 0; kk < 32; ++kk) {
                    const size_t off = (size_t)(4 * kk) * LX_S2 + 16 * t;
                    mz[t][kk] = pz[off];
                    m2[t][kk] = pz[2 * LX_S2_MAT + off];
                    m1[t][kk] = pz[LX_S2_MAT + off];
                }
            // thread: row r, columns cb + 2i; every slot's load issued before any poll
            const int r = tid & (LX_S2 - 1), cb = tid >> 7;
            const int need = lx_s2_nparts(s, C);
            const T *pp = part + (size_t)rowbase[s] * 16 * LX_S2 + r;
            const int ip2 = TRANS ? ib + 2 : ib - 2;
            const bool h2 = s >= 2 && r < min(LX_S2, n - ip2 * LX_S2);
            const T *p2 = mb + ip2 * LX_S2 + r;
            {
                T y[8];
                U xv[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int cc = cb + 2 * i;
                    xv[i] = (h2 && cc < nrhs) ? lx_sw_mb_peek<T>(p2 + cc * ldmb) : U(0);
                    y[i] = (cc < nrhs && r < hb) ? X[(int64_t)cc * ldx + r0 + r] : T(0);
                }
                // the parts in groups of four, oldest group first (the newest part,
                // q = 0, is the one still being computed)
                for (int q0 = (need - 1) / 4 * 4; q0 >= 0; q0 -= 4) {
                    U pv[4][8];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const int cc = cb + 2 * i;
                            pv[q][i] = (q0 + q < need && cc < nrhs)
                                           ? lx_sw_mb_peek<T>(pp + (size_t)((q0 + q) * 16 + cc) * LX_S2)
                                           : U(0);
                        }
#pragma unroll
                    for (int q = 3; q >= 0; --q)
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const int cc = cb + 2 * i;
                            if (q0 + q < need && cc < nrhs)
                                y[i] -= lx_sw_mb_finish<T>(pp + (size_t)((q0 + q) * 16 + cc) * LX_S2, pv[q][i], abort,
                                                           spin_ticks);
                        }
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int cc = cb + 2 * i;
                    sY[cc][r] = y[i];
                    sX2[cc][r] = (h2 && cc < nrhs) ? lx_sw_mb_finish<T>(p2 + cc * ldmb, xv[i], abort, spin_ticks) : T(0);
                }
            }
            __syncthreads();
            // U = Z Y - M2 x_{s-2}: tile t of wave w, value i = row 32w + 16t + 4lk + i, column lr
            typename MF::acc ua[2], ub[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) ua[t][i] = ub[t][i] = T(0);
#pragma unroll
            for (int kk = 0; kk < 32; ++kk) {
                const T by = sY[lr][4 * kk + lk], bx = sX2[lr][4 * kk + lk];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    ua[t] = MF::mac(mz[t][kk], by, ua[t]);
                    ub[t] = MF::mac(m2[t][kk], bx, ub[t]);
                }
            }
            T u[2][4];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) u[t][i] = ua[t][i] - ub[t][i];
            if (s >= 1) {  // uniform: the chain, x = U - M1 x_{s-1}
                const int ip1 = TRANS ? ib + 1 : ib - 1;
                const bool h1 = r < min(LX_S2, n - ip1 * LX_S2);
                const T *p1 = mb + ip1 * LX_S2 + r;
                U e[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int cc = cb + 2 * i;
                    e[i] = (h1 && cc < nrhs) ? lx_sw_mb_peek<T>(p1 + cc * ldmb) : U(0);
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int cc = cb + 2 * i;
                    sX1[cc][r] = (h1 && cc < nrhs) ? lx_sw_mb_finish<T>(p1 + cc * ldmb, e[i], abort, spin_ticks) : T(0);
                }
                __syncthreads();
                typename MF::acc xa[2][2];
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int i = 0; i < 4; ++i) xa[t][h][i] = T(0);
#pragma unroll
                for (int kk = 0; kk < 32; ++kk) {
                    const T bx = sX1[lr][4 * kk + lk];
#pragma unroll
                    for (int t = 0; t < 2; ++t) xa[t][kk & 1] = MF::mac(m1[t][kk], bx, xa[t][kk & 1]);
                }
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i) u[t][i] -= xa[t][0][i] + xa[t][1][i];
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 32 * w + 16 * t + MF::row(lk, i);
                    if (row < hb && lr < nrhs) {
                        X[(int64_t)lr * ldx + r0 + row] = u[t][i];
                        lx_sw_mb_put<T>(mb + lr * ldmb + r0 + row, u[t][i]);
                    }
                }
        } else {
            // ---------------- P(s, c) ----------------
            int t_lo, t_hi;
            lx_s2_part(s, c, C, t_lo, t_hi);
            const int rh = w >> 1, ch = w & 1;
            const int hr = min(64, hb - 64 * rh);
            typename MF::acc macc[4];
#pragma unroll
            for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                for (int i = 0; i < 4; ++i) macc[tt][i] = T(0);
            T(*sx)[68] = sxp[w];
            if (hr > 0) {  // wave-uniform
                const int rbase = r0 + 64 * rh;
                for (int t = t_lo; t < t_hi; ++t) {
                    const int jb = TRANS ? nb - 1 - t : t;
                    const int c0 = jb * LX_S2 + 64 * ch, hj = min(64, n - c0);
                    if (hj <= 0) continue;  // uniform (a ragged last block, backward)
                    // the block of L before the wait: only x depends on step t
                    T a[64];
                    if (!TRANS) {
                        // a[tt*16 + kk] = L[rbase + 16 tt + lr, c0 + 4 kk + lk] (rows past hr clamped)
#pragma unroll
                        for (int tt = 0; tt < 4; ++tt) {
                            const T *Lb = L + (int64_t)(c0 + lk) * ldl + rbase + min(16 * tt + lr, hr - 1);
#pragma unroll
                            for (int kk = 0; kk < 16; ++kk) a[tt * 16 + kk] = Lb[(int64_t)(4 * kk) * ldl];
                        }
                    } else {
                        const int kk = lane % KC, cq = lane / KC;
#pragma unroll
                        for (int m = 0; m < 64 / KC; ++m)
#pragma unroll
                            for (int q = 0; q < 64 / CPI; ++q) {
                                const int colg = rbase + min(q * CPI + cq, hr - 1);
                                const int rowg = c0 + min(m * KC + kk, hj - 1);
                                a[m * (64 / CPI) + q] = L[(int64_t)colg * ldl + rowg];
                            }
                    }
                    // x of the block: lane = row, 16 columns, loads issued before any poll
                    {
                        U e[16];
                        const T *px = mb + c0 + lane;
#pragma unroll
                        for (int j = 0; j < 16; ++j) e[j] = (lane < hj && j < nrhs) ? lx_sw_mb_peek<T>(px + j * ldmb) : U(0);
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            sx[j][lane] = (lane < hj && j < nrhs) ? lx_sw_mb_finish<T>(px + j * ldmb, e[j], abort, spin_ticks) : T(0);
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (!TRANS) {
#pragma unroll
                        for (int kk = 0; kk < 16; ++kk) {
                            const T bv = sx[lr][4 * kk + lk];
#pragma unroll
                            for (int tt = 0; tt < 4; ++tt) macc[tt] = MF::mac(a[tt * 16 + kk], bv, macc[tt]);
                        }
                    } else {
                        const int kk = lane % KC, cq = lane / KC;
#pragma unroll
                        for (int m = 0; m < 64 / KC; ++m) {
#pragma unroll
                            for (int q = 0; q < 64 / CPI; ++q) stile[w][q * CPI + cq][kk] = a[m * (64 / CPI) + q];
                            __builtin_amdgcn_wave_barrier();
#pragma unroll
                            for (int k4 = 0; k4 < KC; k4 += 4) {  // rows past hj: x is zero there
                                const T bv = sx[lr][m * KC + k4 + lk];
#pragma unroll
                                for (int tt = 0; tt < 4; ++tt)
                                    macc[tt] = MF::mac(stile[w][16 * tt + lr][k4 + lk], bv, macc[tt]);
                            }
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // the two column halves of each row half meet in LDS, in a fixed order
#pragma unroll
            for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                for (int i = 0; i < 4; ++i) sred[ch][lr][64 * rh + 16 * tt + MF::row(lk, i)] = macc[tt][i];
            __syncthreads();
            T *pp = part + ((size_t)rowbase[s] + (c - 1)) * 16 * LX_S2;
            const int r = tid & (LX_S2 - 1), cb = tid >> 7;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int cc = cb + 2 * i;
                if (cc < nrhs)
                    lx_sw_mb_put<T>(pp + (size_t)cc * LX_S2 + r, r < hb ? sred[0][cc][r] + sred[1][cc][r] : T(0));
            }
        }
    }
}

#ifndef LX_SWEEP_KERNELS_ONLY
#include <map>
#include <mutex>

namespace {

// 64-row blocks per partial-sum unit: at least 8, and at most 8 parts per step
constexpr int lx_sw_chunk(int nb) { return (nb - 1 + 7) / 8 > 8 ? (nb - 1 + 7) / 8 : 8; }

struct lx_sw_plan {
    int *d_units = nullptr, *d_rowbase = nullptr;
    int nunits = 0, nparts = 0;
};
// Per-device properties: CUs, wall clock, resident workgroups of each sweep
// kernel ([type][trans][NR 1, 4, 16]).  Filled once per device, in a fixed
// table (no allocation), so the scratch-free entry point can use it too.
struct lx_sw_props {
    int cus = 0, wall_khz = 0, occ[2][2][3] = {};
};
constexpr int LX_SW_MAXDEV = 64;
std::mutex lx_sw_mu;
lx_sw_props lx_sw_prop_tab[LX_SW_MAXDEV];
std::map<int, std::map<int, lx_sw_plan>> lx_sw_plans;  // [device][number of blocks]

template <typename T, int TR, int NR>
int lx_sw_occ()
{
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, lx_sweep_kernel<T, TR, NR>, 256, 0) != hipSuccess || o < 1)
        o = 1;
    return o;
}
template <typename T>
void lx_sw_occ_all(int (&o)[2][3])
{
    o[0][0] = lx_sw_occ<T, 0, 1>();
    o[0][1] = lx_sw_occ<T, 0, 4>();
    o[0][2] = lx_sw_occ<T, 0, 16>();
    o[1][0] = lx_sw_occ<T, 1, 1>();
    o[1][1] = lx_sw_occ<T, 1, 4>();
    o[1][2] = lx_sw_occ<T, 1, 16>();
}

// Properties of the current device (caller holds lx_sw_mu).
const lx_sw_props *lx_sw_props_get(int *dev_out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= LX_SW_MAXDEV) return nullptr;
    lx_sw_props &d = lx_sw_prop_tab[dev];
    if (d.cus == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return nullptr;
        if (hipDeviceGetAttribute(&d.wall_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || d.wall_khz <= 0)
            d.wall_khz = 100000;
        lx_sw_occ_all<double>(d.occ[0]);
        lx_sw_occ_all<float>(d.occ[1]);
        d.cus = cus;
    }
    if (dev_out) *dev_out = dev;
    return &d;
}

// Properties and the unit list for nb blocks (built once per device and nb).
const lx_sw_props *lx_sw_get(int nb, const lx_sw_plan **plan)
{
    std::lock_guard<std::mutex> g(lx_sw_mu);
    int dev = 0;
    const lx_sw_props *d = lx_sw_props_get(&dev);
    if (!d) return nullptr;
    auto &plans = lx_sw_plans[dev];
    auto it = plans.find(nb);
    if (it == plans.end()) {
        std::vector<int> units, rowbase;
        lx_sw_build_units(nb, lx_sw_chunk(nb), units, rowbase);
        lx_sw_plan pl;
        pl.nunits = (int)units.size();
        pl.nparts = rowbase[nb];
        if (hipMalloc(&pl.d_units, sizeof(int) * units.size()) != hipSuccess) return nullptr;
        if (hipMalloc(&pl.d_rowbase, sizeof(int) * rowbase.size()) != hipSuccess) {
            (void)hipFree(pl.d_units);
            return nullptr;
        }
        if (hipMemcpy(pl.d_units, units.data(), sizeof(int) * units.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(pl.d_rowbase, rowbase.data(), sizeof(int) * rowbase.size(), hipMemcpyHostToDevice) != hipSuccess)
            return nullptr;
        it = plans.emplace(nb, pl).first;
    }
    *plan = &it->second;
    return d;
}

// Workspace of one sweep launch over nb blocks with nparts partial sums:
// ctl (ticket, abort, flags, part counters) | part | mailbox, each 256-byte
// aligned.  Offsets in bytes for element size es.
struct lx_sw_ws {
    size_t ctl_ints, part_off, mb_off, bytes;
    int64_t ldmb;
};
constexpr lx_sw_ws lx_sw_layout(int nb, int nparts, size_t es)
{
    const size_t ctl_ints = 64 + 2 * (size_t)nb;
    const size_t part_off = (ctl_ints * sizeof(int) + 255) / 256 * 256;
    const size_t part_bytes = es * (size_t)(nparts > 1 ? nparts : 1) * LX_SW_NR * 64;
    const size_t mb_off = (part_off + part_bytes + 255) / 256 * 256;
    const int64_t ldmb = (int64_t)nb * 64;
    return lx_sw_ws{ctl_ints, part_off, mb_off, (mb_off + es * (size_t)ldmb * LX_SW_NR + 255) / 256 * 256, ldmb};
}
}  // namespace

// The sweeps of X := op(L)^-1 X (which: 1 forward, 2 backward, 3 both) over
// the given unit list and workspace, 16 right-hand sides per launch; every
// launch is waited for and checked for a wait that gave up.
template <typename T>
static int sweeps_launch(const lx_sw_props *d, int n, int nrhs, const T *L, int64_t ldl, T *X, int64_t ldx, int which,
                         hipStream_t s, const int *d_units, int nunits, const int *d_rowbase, int nparts, char *ws)
{
    const int nb = (n + 63) / 64;
    const lx_sw_ws lay = lx_sw_layout(nb, nparts, sizeof(T));
    int *ctl = (int *)ws;
    T *part = (T *)(ws + lay.part_off);
    T *mb = (T *)(ws + lay.mb_off);
    const int64_t ldmb = lay.ldmb;
    double tmo = 30.0;
    if (const char *ev = getenv("LINOMPSS_TIMEOUT_S")) tmo = atof(ev) > 0 ? atof(ev) : tmo;
    const long long ticks = (long long)(tmo * d->wall_khz * 1000.0);
    const int ti = sizeof(T) == 8 ? 0 : 1;
    int aborted = 0;
    for (int c0 = 0; c0 < nrhs; c0 += LX_SW_NR) {
        const int nc = std::min(LX_SW_NR, nrhs - c0);
        const int ri = nc <= 1 ? 0 : nc <= 4 ? 1 : 2;
        for (int tr = 0; tr < 2; ++tr) {
            if (!(which & (1 << tr))) continue;
            const int grid = std::max(1, std::min(nunits, d->cus * d->occ[ti][tr][ri]));
            lx_fill_i32(ctl, 0, (int64_t)lay.ctl_ints, s);
            lx_fill_i32((int32_t *)mb, -1, (int64_t)(sizeof(T) * ldmb * nc / 4), s);  // all-ones: empty
#define LX_SW_LAUNCH(TR, NR)                                                                                       \
    hipLaunchKernelGGL((lx_sweep_kernel<T, TR, NR>), dim3(grid), dim3(256), 0, s, n, nc, L, ldl, X + c0 * ldx, ldx,  \
                       d_units, nunits, d_rowbase, lx_sw_chunk(nb), part, ctl, mb, ldmb, ticks)
            if (tr == 0) {
                if (ri == 0) LX_SW_LAUNCH(0, 1);
                else if (ri == 1) LX_SW_LAUNCH(0, 4);
                else LX_SW_LAUNCH(0, 16);
            } else {
                if (ri == 0) LX_SW_LAUNCH(1, 1);
                else if (ri == 1) LX_SW_LAUNCH(1, 4);
                else LX_SW_LAUNCH(1, 16);
            }
#undef LX_SW_LAUNCH
            if (hipGetLastError() != hipSuccess) return LINOMPSS_ERR_HIP;
            int ab = 0;
            LX_HIP_TRY(hipMemcpyAsync(&ab, ctl + 1, sizeof(int), hipMemcpyDeviceToHost, s));
            LX_HIP_TRY(hipStreamSynchronize(s));
            aborted |= ab;
        }
    }
    return aborted ? LINOMPSS_ERR_SCHED : 0;
}

template <typename T>
static int sweeps(int n, int nrhs, const T *L, int64_t ldl, T *X, int64_t ldx, int which, hipStream_t s)
{
    const int nb = (n + 63) / 64;
    const lx_sw_plan *pl = nullptr;
    const lx_sw_props *d = lx_sw_get(nb, &pl);
    if (!d) return LINOMPSS_ERR_HIP;
    char *ws = (char *)lx_scratch(3, lx_sw_layout(nb, pl->nparts, sizeof(T)).bytes);
    if (!ws) return LINOMPSS_ERR_NOMEM;
    return sweeps_launch<T>(d, n, nrhs, L, ldl, X, ldx, which, s, pl->d_units, pl->nunits, pl->d_rowbase, pl->nparts,
                            ws);
}

// which: 1 forward, 2 backward, 3 both.  Returns after the sweeps finished
// (each launch is checked for a timed-out wait).
// ---------------------------------------------------------------------------
// 128-row FP32 sweeps (lx_sweep2_kernel for one right-hand side,
// lx_sweep2m_kernel for up to 16 a launch) over the step matrices of
// lx_sw2_prep_kernel.
namespace {
std::map<int, std::map<int, lx_sw_plan>> lx_s2_plans;  // [device][1024 * blocks + newest part's steps]
int lx_s2_occ_tab[LX_SW_MAXDEV][3][2];  // [device][one column at 1 or 2 waves per SIMD, 16 columns][trans]
}  // namespace

int lx_sw2_enabled(int dtype, int n, int nrhs)
{
    if (dtype != 1 || nrhs < 1) return 0;
    int min_n = 2048;  // below it the pre-pass costs more than the shorter chain saves
    if (const char *e = getenv("LINOMPSS_SWEEP2_MIN_N")) min_n = atoi(e);
    return min_n > 0 && n >= min_n;
}

size_t lx_sw2_prep_bytes(int n) { return sizeof(float) * 6 * (size_t)LX_S2_MAT * (size_t)((n + LX_S2 - 1) / LX_S2); }

int lx_sw2_prep(int n, const float *L, int64_t ldl, float *prep, hipStream_t s)
{
    if (n <= 0) return 0;
    const int nb = (n + LX_S2 - 1) / LX_S2;
    hipLaunchKernelGGL((lx_sw2_prep_kernel<float, 0>), dim3(nb), dim3(256), 0, s, n, L, ldl, prep);
    hipLaunchKernelGGL((lx_sw2_prep_kernel<float, 1>), dim3(nb), dim3(256), 0, s, n, L, ldl, prep);
    return hipGetLastError() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

int lx_sw2_sweeps(int n, int nrhs, const float *L, int64_t ldl, float *X, int64_t ldx, int which, const float *prep,
                  hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return 0;
    const int nb = (n + LX_S2 - 1) / LX_S2;
    const bool multi = nrhs > 1;  // 16 columns a launch on the matrix cores, else one column
    int C = 2;  // steps of the newest partial sum (then 2C, 4C, ...)
    if (const char *e = getenv("LINOMPSS_SWEEP2_PART")) C = std::min(1023, std::max(1, atoi(e)));
    const int W = multi ? 16 : 1;  // columns held by the mailbox and the parts
    const lx_sw_props *d = nullptr;
    lx_sw_plan pl;
    int occ[2];
    int wpe = 1;  // LINOMPSS_SWEEP2_WPE=2: the one-column build budgeted for two waves per SIMD
    if (const char *e = getenv("LINOMPSS_SWEEP2_WPE")) wpe = atoi(e) == 2 ? 2 : 1;
    {
        std::lock_guard<std::mutex> g(lx_sw_mu);
        int dev = 0;
        d = lx_sw_props_get(&dev);
        if (!d) return LINOMPSS_ERR_HIP;
        if (lx_s2_occ_tab[dev][0][0] == 0) {
            auto occ_of = [](const void *k) {
                int o = 0;
                const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k, 256, 0);
                return (e == hipSuccess && o >= 1) ? o : 1;
            };
            lx_s2_occ_tab[dev][0][0] = occ_of((const void *)lx_sweep2_kernel<float, 0, 1>);
            lx_s2_occ_tab[dev][0][1] = occ_of((const void *)lx_sweep2_kernel<float, 1, 1>);
            lx_s2_occ_tab[dev][1][0] = occ_of((const void *)lx_sweep2_kernel<float, 0, 2>);
            lx_s2_occ_tab[dev][1][1] = occ_of((const void *)lx_sweep2_kernel<float, 1, 2>);
            lx_s2_occ_tab[dev][2][0] = occ_of((const void *)lx_sweep2m_kernel<float, 0>);
            lx_s2_occ_tab[dev][2][1] = occ_of((const void *)lx_sweep2m_kernel<float, 1>);
        }
        const int oi = multi ? 2 : wpe - 1;
        occ[0] = lx_s2_occ_tab[dev][oi][0];
        occ[1] = lx_s2_occ_tab[dev][oi][1];
        auto &plans = lx_s2_plans[dev];
        const int key = nb * 1024 + C;
        auto it = plans.find(key);
        if (it == plans.end()) {
            std::vector<int> units, rowbase;
            lx_s2_build_units(nb, C, units, rowbase);
            lx_sw_plan np;
            np.nunits = (int)units.size();
            np.nparts = rowbase[nb];
            if (hipMalloc(&np.d_units, sizeof(int) * units.size()) != hipSuccess) return LINOMPSS_ERR_NOMEM;
            if (hipMalloc(&np.d_rowbase, sizeof(int) * rowbase.size()) != hipSuccess) {
                (void)hipFree(np.d_units);
                return LINOMPSS_ERR_NOMEM;
            }
            if (hipMemcpy(np.d_units, units.data(), sizeof(int) * units.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(np.d_rowbase, rowbase.data(), sizeof(int) * rowbase.size(), hipMemcpyHostToDevice) !=
                    hipSuccess)
                return LINOMPSS_ERR_HIP;
            it = plans.emplace(key, np).first;
        }
        pl = it->second;
    }
    // workspace: ctl (ticket, abort) | part | mailbox
    const size_t ctl_ints = 64;
    const size_t part_off = (ctl_ints * sizeof(int) + 255) / 256 * 256;
    const size_t mb_off = (part_off + sizeof(float) * (size_t)std::max(1, pl.nparts) * LX_S2 * W + 255) / 256 * 256;
    const int64_t ldmb = (int64_t)nb * LX_S2;
    const size_t bytes = mb_off + sizeof(float) * (size_t)ldmb * W;
    char *ws = (char *)lx_scratch(3, bytes);
    if (!ws) return LINOMPSS_ERR_NOMEM;
    int *ctl = (int *)ws;
    float *part = (float *)(ws + part_off), *mb = (float *)(ws + mb_off);
    double tmo = 30.0;
    if (const char *ev = getenv("LINOMPSS_TIMEOUT_S")) tmo = atof(ev) > 0 ? atof(ev) : tmo;
    const long long ticks = (long long)(tmo * d->wall_khz * 1000.0);
    int aborted = 0;
    for (int c0 = 0; c0 < nrhs; c0 += 16) {
        const int nc = multi ? std::min(16, nrhs - c0) : 1;
        float *Xc = X + (int64_t)c0 * ldx;
        for (int tr = 0; tr < 2; ++tr) {
            if (!(which & (1 << tr))) continue;
            const int grid = std::max(1, std::min(pl.nunits, d->cus * occ[tr]));
            lx_fill_i32(ctl, 0, (int64_t)ctl_ints, s);
            // all-ones: empty (the parts and the mailbox are contiguous)
            lx_fill_i32((int32_t *)part, -1, (int64_t)((mb_off - part_off) / sizeof(float)) + ldmb * W, s);
#define LX_S2_LAUNCH(TR, WPE)                                                                                      \
    hipLaunchKernelGGL((lx_sweep2_kernel<float, TR, WPE>), dim3(grid), dim3(256), 0, s, n, L, ldl, Xc, prep,      \
                       pl.d_units, pl.nunits, pl.d_rowbase, C, part, ctl, mb, ticks)
#define LX_S2M_LAUNCH(TR)                                                                                          \
    hipLaunchKernelGGL((lx_sweep2m_kernel<float, TR>), dim3(grid), dim3(256), 0, s, n, nc, L, ldl, Xc, ldx, prep, \
                       pl.d_units, pl.nunits, pl.d_rowbase, C, part, ctl, mb, ldmb, ticks)
            if (multi) {
                if (tr == 0) LX_S2M_LAUNCH(0);
                else LX_S2M_LAUNCH(1);
            } else if (tr == 0) {
                if (wpe == 2) LX_S2_LAUNCH(0, 2);
                else LX_S2_LAUNCH(0, 1);
            } else {
                if (wpe == 2) LX_S2_LAUNCH(1, 2);
                else LX_S2_LAUNCH(1, 1);
            }
#undef LX_S2_LAUNCH
#undef LX_S2M_LAUNCH
            if (hipGetLastError() != hipSuccess) return LINOMPSS_ERR_HIP;
            int ab = 0;
            LX_HIP_TRY(hipMemcpyAsync(&ab, ctl + 1, sizeof(int), hipMemcpyDeviceToHost, s));
            LX_HIP_TRY(hipStreamSynchronize(s));
            aborted |= ab;
        }
    }
    return aborted ? LINOMPSS_ERR_SCHED : 0;
}

int lx_sweeps_dev(int dtype, int n, int nrhs, const void *L, int64_t ldl, void *X, int64_t ldx, int which,
                  hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return 0;
    if (dtype == 0) return sweeps<double>(n, nrhs, (const double *)L, ldl, (double *)X, ldx, which, s);
    if (lx_sw2_enabled(dtype, n, nrhs)) {
        float *prep = (float *)lx_scratch(13, lx_sw2_prep_bytes(n));
        if (!prep) return LINOMPSS_ERR_NOMEM;
        int rc = lx_sw2_prep(n, (const float *)L, ldl, prep, s);
        if (rc) return rc;
        return lx_sw2_sweeps(n, nrhs, (const float *)L, ldl, (float *)X, ldx, which, prep, s);
    }
    return sweeps<float>(n, nrhs, (const float *)L, ldl, (float *)X, ldx, which, s);
}

// ---------------------------------------------------------------------------
// Scratch-free solve with an existing factor (linompss_potrs_dev_noalloc).
//
// For callers that cannot allocate device memory (code under stream capture,
// serving loops, harnesses that forbid hipMalloc): the sweeps' workspace and
// their unit list live in one file-scope device array, sized for n <= 32768
// and either element type, that each device holds until the process ends.
// Host code never passes the array's name to a kernel: on the first call on a
// device the array's device address is looked up with hipGetSymbolAddress
// (with that device current) and kept per device.  Every call on a device
// shares the array and overwrites it (counters and mailbox are cleared by
// fill kernels at each launch, the unit list is rebuilt by a one-thread
// kernel at each call), so calls that overlap on one device race; within
// this library they are serialised by the device lock.
#define LX_NA_MAX_N 32768
namespace {
// partial sums of a sweep over nb blocks: sum over s = 4 .. nb-1 of
// ceil((s - 3) / C), in closed form (q full groups of C, then r more)
constexpr int lx_na_nparts(int nb)
{
    const int C = lx_sw_chunk(nb), m = nb - 1 - LX_SW_GAP;
    if (m <= 0) return 0;
    const int q = m / C, r = m % C;
    return C * q * (q + 1) / 2 + r * (q + 1);
}
// bytes of the workspace plus the unit list (nb + nparts codes) and the
// part offsets (nb + 1), for nb blocks
constexpr size_t lx_na_bytes(int nb)
{
    const int np = lx_na_nparts(nb);
    return lx_sw_layout(nb, np, sizeof(double)).bytes + ((size_t)(nb + np) + (size_t)(nb + 1)) * sizeof(int) + 256;
}
constexpr size_t lx_na_max_bytes()
{
    size_t m = 0;
    for (int nb = 1; nb <= LX_NA_MAX_N / 64; ++nb) m = lx_na_bytes(nb) > m ? lx_na_bytes(nb) : m;
    return m;
}
constexpr size_t LX_NA_BYTES = (lx_na_max_bytes() + 255) / 256 * 256;
static_assert(LX_NA_BYTES <= (size_t)24 << 20, "the scratch-free solve keeps its array under 24 MiB");
}  // namespace

static __device__ __align__(256) char scratch[LX_NA_BYTES];

// The unit list of lx_sw_build_units, written by one thread on the device.
__global__ void lx_sw_units_kernel(int nb, int C, int *__restrict__ units, int *__restrict__ rowbase)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int acc = 0;
    rowbase[0] = 0;
    for (int s2 = 0; s2 < nb; ++s2) {
        acc += lx_sw_nparts(s2, C);
        rowbase[s2 + 1] = acc;
    }
    int u = 0;
    for (int key = -1; key <= nb - 2; ++key) {
        units[u++] = (key + 1) << 16;
        if (key < 0) continue;
        for (int c = 1;; ++c) {
            const long long st = (long long)key + LX_SW_GAP + 1 + (long long)(c - 1) * C;
            if (st >= nb) break;
            units[u++] = (int)(st << 16) | c;
        }
    }
}

static void *lx_na_ptr[LX_SW_MAXDEV];

extern "C" int linompss_potrs_dev_noalloc(int dtype, int n, int nrhs, int b, const void *dL, int64_t lda, void *dX,
                                          int64_t ldx, void *stream, const linompss_exec_opts *opts)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (b <= 0) return -4;
    if (lda < std::max(1, n)) return -6;
    if (ldx < std::max(1, n)) return -8;
    if (opts && opts->fused_solve) return -10;  // the fused form runs on the executor's own workspace
    if (n > LX_NA_MAX_N) return LINOMPSS_ERR_LIMIT;
    if (n == 0 || nrhs == 0) return 0;
    std::lock_guard<std::mutex> g(lx_dev_mutex());
    std::lock_guard<std::mutex> g2(lx_sw_mu);
    int dev = 0;
    const lx_sw_props *d = lx_sw_props_get(&dev);
    if (!d) return LINOMPSS_ERR_NODEV;
    if (!lx_na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LINOMPSS_ERR_HIP;
        lx_na_ptr[dev] = p;
    }
    char *base = (char *)lx_na_ptr[dev];
    const int nb = (n + 63) / 64, C = lx_sw_chunk(nb), np = lx_na_nparts(nb);
    const size_t es = dtype == 0 ? sizeof(double) : sizeof(float);
    const lx_sw_ws lay = lx_sw_layout(nb, np, es);
    int *units = (int *)(base + lay.bytes);
    int *rowbase = units + (nb + np);
    if ((size_t)((char *)(rowbase + nb + 1) - base) > LX_NA_BYTES) return LINOMPSS_ERR_LIMIT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lx_sw_units_kernel, dim3(1), dim3(64), 0, s, nb, C, units, rowbase);
    if (hipGetLastError() != hipSuccess) return LINOMPSS_ERR_HIP;
    int rc = dtype == 0 ? sweeps_launch<double>(d, n, nrhs, (const double *)dL, lda, (double *)dX, ldx, 3, s, units,
                                                nb + np, rowbase, np, base)
                        : sweeps_launch<float>(d, n, nrhs, (const float *)dL, lda, (float *)dX, ldx, 3, s, units,
                                               nb + np, rowbase, np, base);
    return rc;
}
#endif  // LX_SWEEP_KERNELS_ONLY
