// lx_exec_kernel.hpp -- the persistent tile-DAG executor kernel (device
// code only; the host side lives in lx_exec.hip).
//
// Every workgroup loops: grab the next unit index (one atomic), wait until
// each of the unit's dependency counters reaches its threshold (relaxed
// polling, then one acquire fence), run the tile kernel, make its stores
// visible (release fence in every wave), and bump the unit's tile counter.
// Units are listed in a topological order chosen by the host planner, so
// every wait is on a unit that a running workgroup already holds: the launch
// cannot deadlock, whatever the number of resident workgroups.  A wait that
// exceeds spin_ticks sets an abort flag that every workgroup observes, so each
// wave reaches the kernel's exit.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>

#include "lx_device.hpp"
#include "lx_plan.hpp"

// Op families compiled into one executor kernel.  Each family is its own
// kernel symbol, so a launch carries only the registers and LDS its units use.
enum lx_family : int { LX_FAM_FACTOR = 0, LX_FAM_SOLVE = 1, LX_FAM_ALL = 2 };
template <typename T, int FAM> struct lx_fam_smem {
    // POTRF (2 * 64 * 65 * sizeof(T) + 16; its trailing update reuses the space
    // for the staged GEMM), TRSM (Z_q, then its update slabs or Y), staged GEMM
    static constexpr int potrf = 2 * 64 * 65 * (int)sizeof(T) + 16;
    static constexpr int gemm0 = lx_lgemm_smem<T>::bytes > lx_bx_smem::bytes ? lx_lgemm_smem<T>::bytes : lx_bx_smem::bytes;
    static constexpr int gemm1 = gemm0 > lx_bxw_smem::bytes ? gemm0 : lx_bxw_smem::bytes;
    static constexpr int gemm = gemm1 > lx_p16_smem::bytes ? gemm1 : lx_p16_smem::bytes;
    static constexpr int trsm = 64 * 65 * (int)sizeof(T) + (4 * 16 * 65 * (int)sizeof(T) > lx_trsm_upd_smem<T>::bytes
                                                                 ? 4 * 16 * 65 * (int)sizeof(T)
                                                                 : lx_trsm_upd_smem<T>::bytes);
    static constexpr int factor0 = potrf > gemm ? potrf : gemm;
    static constexpr int factor = factor0 > trsm ? factor0 : trsm;
    static constexpr int solve = 42 * 1024;
    static constexpr int bytes = FAM == LX_FAM_FACTOR ? factor : FAM == LX_FAM_SOLVE ? solve : (factor > solve ? factor : solve);
};
template <int FAM> struct lx_fam_occ { static constexpr int waves_per_simd = FAM == LX_FAM_FACTOR ? 2 : 1; };

template <typename T> struct lx_gemm_cfg;
template <> struct lx_gemm_cfg<double> { static constexpr int FM = 4, FN = 4, U = 2; };
template <> struct lx_gemm_cfg<float> { static constexpr int FM = 4, FN = 4, U = 4; };
static constexpr int LX_UNIT_M = 128, LX_UNIT_N = 128;  // 2 x 2 waves of 16*FM x 16*FN

// The columns of a split-body GEMM / SYRK unit: 128 x 256 blocks while more
// than 128 columns remain (aux0 bit 2), then 128 x 128.
template <bool H16>
__device__ __forceinline__ void lx_split_blocks(const lx_unit &u, float *C, long long ldc, const float *Ap,
                                                long long lda, const float *Bp, long long ldb, bool lower, char *smem)
{
    const bool wide = (u.aux0 & 4) != 0;
    for (int c0 = 0; c0 < u.N;) {  // uniform
        const int rem = u.N - c0;
        if (wide && rem > LX_UNIT_N) {
            const int nc = rem < 2 * LX_UNIT_N ? rem : 2 * LX_UNIT_N;
            lx_unit_gemm_bx6w<H16>(C + (long long)c0 * ldc, ldc, Ap, lda, Bp + c0, ldb, u.M, nc, u.K, lower,
                                   u.aux1 - c0, smem);
            c0 += nc;
        } else {
            const int nc = rem < LX_UNIT_N ? rem : LX_UNIT_N;
            lx_unit_gemm_bx6<H16>(C + (long long)c0 * ldc, ldc, Ap, lda, Bp + c0, ldb, u.M, nc, u.K, lower,
                                  u.aux1 - c0, smem);
            c0 += nc;
        }
    }
}

__device__ __forceinline__ long long lx_off(int64_t rcp, long long ld)
{
    return (long long)(rcp >> 32) * ld + (long long)(rcp & 0xffffffffll);
}

// The trailing update of rows r0 .. r0 + M - 1 of a GEMM / SYRK unit (C, A
// and the unit's global offsets advanced by r0; M <= 128).
template <typename T, bool HY>
__device__ __forceinline__ void lx_run_update_rows(const lx_unit &u0, const lx_sys &s, T *C, long long ldc,
                                                   const T *Ap, long long lda, const T *Bp, long long ldb, int r0,
                                                   int M, char *smem)
{
    lx_unit u = u0;
    u.M = M;
    u.aux1 += r0;
    u.a += r0;
    const bool lower = (u.aux0 & 1) != 0;
    if constexpr (sizeof(T) == 4) {
        // split-FP16 units whose operands are whole 16-column slabs of
        // one panel read the planes the TRSM units wrote
        const int ca = (int)(u.a >> 32);
        if (s.planes && (u.aux0 & 10) == 10 && u.K % 16 == 0 && (ca & 15) == 0 && ca == (int)(u.b >> 32)) {
            const int ra = (int)(u.a & 0xffffffffll), rb = (int)(u.b & 0xffffffffll);
            for (int c0 = 0; c0 < u.N; c0 += 2 * LX_UNIT_N) {  // uniform
                const int nc = u.N - c0 < 2 * LX_UNIT_N ? u.N - c0 : 2 * LX_UNIT_N;
                lx_unit_gemm_p16(C + (long long)c0 * ldc, ldc, (const _Float16 *)s.planes, s.ldp, s.nsig, ra,
                                 rb + c0, ca >> 4, u.M, nc, u.K, lower, u.aux1 - c0, smem);
            }
            return;
        }
        if (u.aux0 & 2) {
            if (u.aux0 & 8)
                lx_split_blocks<true>(u, C, ldc, Ap, lda, Bp, ldb, lower, smem);
            else
                lx_split_blocks<false>(u, C, ldc, Ap, lda, Bp, ldb, lower, smem);
            return;
        }
    }
    for (int c0 = 0; c0 < u.N; c0 += LX_UNIT_N) {  // uniform
        const int nc = u.N - c0 < LX_UNIT_N ? u.N - c0 : LX_UNIT_N;
        lx_unit_gemm_lds<T>(C + (long long)c0 * ldc, ldc, Ap, lda, Bp + c0, ldb, u.M, nc, u.K, lower,
                            u.aux1 - c0, smem);
    }
}

template <typename T, int FAM, bool HY = false>
__device__ __forceinline__ void lx_run_unit(const lx_unit &u, const lx_sys *__restrict__ sys, int *info, char *smem)
{
    const lx_sys s = sys[u.sys];
    T *A = (T *)s.A, *X = (T *)s.X;
    T *C;
    const T *Ap, *Bp;
    long long ldc, lda, ldb;
    if constexpr (HY) {
        // hypermatrix (factor plans): a unit's operands never straddle a tile
        // (checked when the plan is built), so each is addressed inside its
        // own tile, leading dimension tb; an unused operand's offset is 0,
        // tile (0, 0)
        auto at = [&](int64_t rcp) {
            const int r = (int)(rcp & 0xffffffffll), c = (int)(rcp >> 32);
            return (T *)s.tiles[(int64_t)(c / s.tb) * s.tmt + r / s.tb] + (int64_t)(c % s.tb) * s.tb + r % s.tb;
        };
        (void)A;
        (void)X;
        ldc = lda = ldb = s.tb;
        C = at(u.c);
        Ap = at(u.a);
        Bp = at(u.b);
    } else {
        T *Cb = u.cbase == LX_BASE_X ? X : A;
        const T *Ab = u.abase == LX_BASE_X ? X : A;
        const T *Bb = u.bbase == LX_BASE_X ? X : A;
        ldc = u.cbase == LX_BASE_X ? s.ldx : s.lda;
        lda = u.abase == LX_BASE_X ? s.ldx : s.lda;
        ldb = u.bbase == LX_BASE_X ? s.ldx : s.lda;
        C = Cb + lx_off(u.c, ldc);
        Ap = Ab + lx_off(u.a, lda);
        Bp = Bb + lx_off(u.b, ldb);
    }
    constexpr bool F = FAM != LX_FAM_SOLVE, S = FAM != LX_FAM_FACTOR;
    switch (u.op) {
    case LX_OP_POTRF: {
        if constexpr (!F) break;
        int loc = 0;
        lx_unit_potrf<T>(C, ldc, u.M, (T *)s.W + u.aux1, s.ldw, &loc, smem);
        if (loc != 0)  // loc is the same in every thread
            atomicMin(&info[s.info_slot], threadIdx.x == 0 ? u.aux1 + loc : INT_MAX);
        break;
    }
    case LX_OP_TRSM_RLTN:
        // the diagonal tile's inverse blocks start at W row = its row in A; a
        // unit of more than 64 rows runs as 64-row passes
        if constexpr (F)
            for (int r0 = 0; r0 < u.M; r0 += 64)  // uniform
                lx_unit_trsm_rltn<T>(C + r0, ldc, Ap, lda, (const T *)s.W + (u.a & 0xffffffffll), s.ldw,
                                     u.M - r0 < 64 ? u.M - r0 : 64, u.N, smem, (_Float16 *)s.planes, s.ldp, s.nsig,
                                     (int)(u.c & 0xffffffffll) + r0, (int)(u.c >> 32));
        break;
    case LX_OP_SYRK:
    case LX_OP_GEMM_NT:
        // a unit covers up to 128 m' rows and N = 128 m columns, done as m' x m
        // staged 128 x 128 blocks (aux1 = global row - global col of C(0, 0));
        // aux0 bit 0: lower mask (SYRK), bit 1: split matrix-core bodies (FP32
        // only), bit 2: split bodies in 128 x 256 blocks while more than 128
        // columns remain, bit 3: the FP16 split instead of the BF16 one
        if constexpr (F) {
            for (int r0 = 0; r0 < u.M; r0 += LX_UNIT_M)  // uniform
                lx_run_update_rows<T, HY>(u, s, C + r0, ldc, Ap + r0, lda, Bp, ldb, r0,
                                          u.M - r0 < LX_UNIT_M ? u.M - r0 : LX_UNIT_M, smem);
        }
        break;
    case LX_OP_TRSV:
        if constexpr (S) lx_unit_trsv_block<T>(C, ldc, Ap, lda, u.M, u.N, u.aux0, smem);
        break;
    case LX_OP_GEMV_RHS:
        if constexpr (S) lx_unit_gemv_rhs<T>(C, ldc, Ap, lda, Bp, ldb, u.M, u.K, u.N, u.aux0, smem);
        break;
    default:
        break;
    }
}

// Wave 0 of the workgroup, all 64 lanes active, grabs the next unit and waits
// for its dependencies; every value that steers its control flow is made
// wave-uniform with readfirstlane.  No branch in this loop separates the lanes
// of one wave: code of the form "if (threadIdx.x == 0) { grab; publish to
// LDS }" lets the compiler send lane 0 and lanes 1..63 of wave 0 around
// different loop structures, with lanes 1..63 waiting at the next barrier for
// a value only lane 0 would write -- a deadlock, since the lanes of a wave
// share one program counter.
__device__ __forceinline__ int lx_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Publishes a unit: counter += 1 once everything the workgroup stored is
// visible device-wide.  Every wave waits for its own stores to reach the L2
// (workgroup-scope release plus an explicit wait: the fence alone emits no
// wait before the barrier), then wave 0 alone issues the agent-scope release,
// whose single L2 write-back covers the whole workgroup, and adds.
// (An agent-scope release fence in each of the four waves costs four L2
// write-backs per unit, measured at tens of microseconds.)
__device__ __forceinline__ void lx_wg_publish(int *counter)
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
        __hip_atomic_fetch_add(counter, (threadIdx.x & 63) == 0 ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Optional profile (prof != null): per workgroup, wave 0 accumulates
// wall-clock ticks spent waiting for dependencies and running each op type,
// written at exit as LX_PROF_WORDS int64 words per workgroup:
//   [0] wait ticks, [1 + op] run ticks (body and publication),
//   [1 + LX_OP_NOP + op] unit counts, [1 + 2 LX_OP_NOP + op] publication ticks.
#define LX_PROF_WORDS (1 + 3 * LX_OP_NOP)

template <typename T, int FAM, bool HY = false>
__global__ void __launch_bounds__(LX_WG_THREADS, lx_fam_occ<FAM>::waves_per_simd)
    lx_exec_kernel(const lx_unit *__restrict__ units, int nunits, const lx_sys *__restrict__ sys, int *ctr,
                   lx_status *st, int *info, long long spin_ticks, long long *prof)
{
    __shared__ __attribute__((aligned(16))) char smem[lx_fam_smem<T, FAM>::bytes];
    __shared__ int s_sched[2];  // [0] = go flag, [1] = unit index
    const int lane = lx_lane();
    __shared__ long long s_prof[LX_PROF_WORDS];  // in LDS: keeps the profile out of the registers
    if (threadIdx.x < LX_PROF_WORDS) s_prof[threadIdx.x] = 0;  // ordered by the loop's first barrier
    long long t_grab = 0, t_ready = 0, t_body = 0;
    for (;;) {
        if (prof) t_grab = wall_clock64();
        if (lx_wave() == 0) {  // wave-uniform branch
            // one atomic per wave: lane 0 adds 1, the others add 0; lane 0's old value is the index
            const int idx = lx_uniform(atomicAdd(&st->head, lane == 0 ? 1 : 0));
            int go = idx < nunits ? 1 : 0;
            if (go && lx_uniform(__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
                go = 0;
            if (go) {
                const int nd = units[idx].nd;
                for (int d = 0; d < nd && go; ++d) {
                    const int c = units[idx].dc[d], thr = units[idx].dt[d];
                    if (lx_uniform(__hip_atomic_load(&ctr[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= thr)
                        continue;
                    const long long t0 = wall_clock64();
                    for (;;) {
                        if (lx_uniform(__hip_atomic_load(&ctr[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= thr)
                            break;
                        if (lx_uniform(__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                            go = 0;
                            break;
                        }
                        // the clock is a scalar read (one value per wave); taken
                        // through readfirstlane so no lane can decide apart
                        if (lx_uniform(wall_clock64() - t0 > spin_ticks ? 1 : 0)) {
                            // every lane stores the same values (no lane branch)
                            atomicAdd(&st->timeouts, lane == 0 ? 1 : 0);
                            atomicOr(&st->abort, lane == 0 ? 1 : 0);  // read-modify-write: seen at once
                            go = 0;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            // every lane of wave 0 stores the same (uniform) values: no lane branch
            s_sched[0] = go;
            s_sched[1] = idx;
        }
        __syncthreads();
        // the same for every lane; readfirstlane makes the loop and the unit's
        // fields wave-uniform (scalar loads, scalar branches around barriers)
        const int go = lx_uniform(s_sched[0]);
        const int idx = lx_uniform(s_sched[1]);
        __syncthreads();  // s_sched is rewritten by wave 0 in the next iteration
        if (!go) break;
        if (prof) t_ready = wall_clock64();
        const lx_unit u = units[idx];
        lx_run_unit<T, FAM, HY>(u, sys, info, smem);
        if (prof) t_body = wall_clock64();
        lx_wg_publish(&ctr[u.ctr]);
        if (lx_wave() == 0) atomicAdd(&st->done, lane == 0 ? 1 : 0);
        if (prof && threadIdx.x == 0 && u.op >= 0 && u.op < LX_OP_NOP) {  // no barrier inside
            const long long t_done = wall_clock64();
            s_prof[0] += t_ready - t_grab;
            s_prof[1 + u.op] += t_done - t_ready;
            s_prof[1 + LX_OP_NOP + u.op] += 1;
            s_prof[1 + 2 * LX_OP_NOP + u.op] += t_done - t_body;
        }
    }
    if (prof && threadIdx.x == 0)
        for (int q = 0; q < LX_PROF_WORDS; ++q) prof[(long long)blockIdx.x * LX_PROF_WORDS + q] = s_prof[q];
}
