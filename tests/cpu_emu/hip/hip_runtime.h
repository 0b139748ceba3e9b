// CPU emulation of the HIP device environment used by the library's kernels
// (test infrastructure: lets the -m "not gpu" tests execute the exact kernel
// source on the host, one workgroup of real OS threads at a time).
//
// * threadIdx / blockIdx are thread_local; one workgroup runs at a time, so
//   __shared__ maps to static storage.
// * __syncthreads is a counting barrier with a timeout: a barrier that not
//   every thread reaches aborts the run with a diagnostic instead of hanging.
// * Lanes run as independent threads, so the emulation checks the kernels'
//   indexing, barrier matching and dependency waits, not how the compiler
//   arranges the lanes of a wave around barriers (only the ISA shows that).
// * Wavefront collectives (MFMA 16x16x4, __shfl, __shfl_xor, readfirstlane) exchange lane
//   values through a per-wave buffer between two wave barriers; a collective
//   that only part of a wave reaches is reported the same way.
// * MFMA operand/result layout: lane l supplies A[l%16][l/16], B[l/16][l%16];
//   lane l receives column l%16, rows 4*(l/16)+i (f32) or (l/16)+4*i (f64).
#pragma once
#include <linux/futex.h>
#include <sys/syscall.h>
#include <time.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>

#define __device__
#define __host__
#define __global__
// kernel attributes of the register allocator have no meaning here
#define LX_WAVES_PER_EU(w)
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

struct lx_emu_dim3 { unsigned x = 0, y = 0, z = 0; };
typedef struct lx_emu_stream *hipStream_t;  // host-side declarations only
extern thread_local lx_emu_dim3 threadIdx;
extern thread_local lx_emu_dim3 blockIdx;
extern lx_emu_dim3 blockDim, gridDim;

// A generation barrier on a futex: arrivals count up, the last one resets the
// count, advances the generation and wakes the sleepers, which re-check the
// generation themselves (no mutex to re-take, so a release does not queue the
// 255 other lanes of a workgroup behind one lock).
struct lx_emu_barrier {
    std::atomic<int> count{0}, gen{0};
    int n = 0;
    const char *name = "";
    void init(int parties, const char *nm)
    {
        n = parties;
        count.store(0);
        gen.store(0);
        name = nm;
    }
    void wait()
    {
        const int g = gen.load(std::memory_order_acquire);
        if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
            count.store(0, std::memory_order_relaxed);
            gen.fetch_add(1, std::memory_order_release);
            syscall(SYS_futex, reinterpret_cast<int *>(&gen), FUTEX_WAKE_PRIVATE, INT_MAX, nullptr, nullptr, 0);
            return;
        }
        const auto t0 = std::chrono::steady_clock::now();
        while (gen.load(std::memory_order_acquire) == g) {
            timespec ts{0, 200000000};  // re-check the deadline every 0.2 s
            syscall(SYS_futex, reinterpret_cast<int *>(&gen), FUTEX_WAIT_PRIVATE, g, &ts, nullptr, 0);
            if (gen.load(std::memory_order_acquire) == g && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
                fprintf(stderr, "lx_emu: %s barrier timeout (%d of %d arrived) -- mismatched barrier\n", name,
                        count.load(), n);
                fflush(stderr);
                std::_Exit(86);
            }
        }
    }
};
extern lx_emu_barrier lx_emu_wg_barrier;
extern lx_emu_barrier lx_emu_wave_barrier[4];
struct lx_emu_wave_buf {
    double a[64], b[64], c[64][4];
};
extern lx_emu_wave_buf lx_emu_wbuf[4];

inline void __syncthreads() { lx_emu_wg_barrier.wait(); }

typedef double lx_emu_d4 __attribute__((ext_vector_type(4)));
typedef float lx_emu_f4 __attribute__((ext_vector_type(4)));

template <typename V, typename S>
inline V lx_emu_mfma(S a, S b, V c)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    lx_emu_wave_buf &B = lx_emu_wbuf[w];
    B.a[lane] = a;
    B.b[lane] = b;
    lx_emu_wave_barrier[w].wait();
    V d = c;
    const int col = lane & 15;
    for (int i = 0; i < 4; ++i) {
        // result rows as the gfx950 builtins deliver them (see lx_device.hpp)
        const int row = sizeof(S) == 8 ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
        S acc = 0;
        for (int k = 0; k < 4; ++k) acc += (S)B.a[k * 16 + row] * (S)B.b[k * 16 + col];
        d[i] = c[i] + acc;
    }
    lx_emu_wave_barrier[w].wait();
    return d;
}
inline lx_emu_d4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, lx_emu_d4 c, int, int, int)
{
    return lx_emu_mfma<lx_emu_d4, double>(a, b, c);
}
inline lx_emu_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, lx_emu_f4 c, int, int, int)
{
    return lx_emu_mfma<lx_emu_f4, float>(a, b, c);
}

// v_mfma_f32_32x32x16_bf16 with the measured gfx950 layout: lane l supplies
// A[l%32][8(l/32)+e] and B[8(l/32)+e][l%32]; result value i of lane l is row
// 8(i/4) + 4(l/32) + i%4, column l%32.  Products are exact in FP32 (8-bit
// significands), summed in FP64 then rounded, close to the hardware's
// accumulation.
typedef __bf16 lx_emu_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lx_emu_f16v __attribute__((ext_vector_type(16)));
struct lx_emu_bf_buf {
    float a[64][8], b[64][8];
};
extern lx_emu_bf_buf lx_emu_bfbuf[4];
inline lx_emu_f16v __builtin_amdgcn_mfma_f32_32x32x16_bf16(lx_emu_bf16x8 a, lx_emu_bf16x8 b, lx_emu_f16v c, int, int,
                                                           int)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    lx_emu_bf_buf &B = lx_emu_bfbuf[w];
    for (int e = 0; e < 8; ++e) {
        B.a[lane][e] = (float)a[e];
        B.b[lane][e] = (float)b[e];
    }
    lx_emu_wave_barrier[w].wait();
    lx_emu_f16v d = c;
    const int col = lane & 31;
    for (int i = 0; i < 16; ++i) {
        const int row = 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
        double acc = 0;
        for (int k = 0; k < 16; ++k) acc += (double)B.a[(k / 8) * 32 + row][k % 8] * (double)B.b[(k / 8) * 32 + col][k % 8];
        d[i] = (float)((double)c[i] + acc);
    }
    lx_emu_wave_barrier[w].wait();
    return d;
}

// v_mfma_f32_32x32x16_f16: the same layout as the BF16 form; products of two
// FP16 numbers (11-bit significands) are exact in FP32.
typedef _Float16 lx_emu_f16x8 __attribute__((ext_vector_type(8)));
inline lx_emu_f16v __builtin_amdgcn_mfma_f32_32x32x16_f16(lx_emu_f16x8 a, lx_emu_f16x8 b, lx_emu_f16v c, int, int, int)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    lx_emu_bf_buf &B = lx_emu_bfbuf[w];
    for (int e = 0; e < 8; ++e) {
        B.a[lane][e] = (float)a[e];
        B.b[lane][e] = (float)b[e];
    }
    lx_emu_wave_barrier[w].wait();
    lx_emu_f16v d = c;
    const int col = lane & 31;
    for (int i = 0; i < 16; ++i) {
        const int row = 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
        double acc = 0;
        for (int k = 0; k < 16; ++k) acc += (double)B.a[(k / 8) * 32 + row][k % 8] * (double)B.b[(k / 8) * 32 + col][k % 8];
        d[i] = (float)((double)c[i] + acc);
    }
    lx_emu_wave_barrier[w].wait();
    return d;
}

template <typename T>
inline T lx_emu_xchg(T v, int src_lane)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    lx_emu_wave_buf &B = lx_emu_wbuf[w];
    B.a[lane] = (double)v;
    lx_emu_wave_barrier[w].wait();
    T r = (T)B.a[src_lane & 63];
    lx_emu_wave_barrier[w].wait();
    return r;
}
template <typename T> inline T __shfl(T v, int src) { return lx_emu_xchg(v, src); }
// readfirstlane: every lane of the wave receives lane 0's value (the emulated
// waves always run with all 64 lanes active)
inline int __builtin_amdgcn_readfirstlane(int v) { return lx_emu_xchg(v, 0); }
inline int __builtin_amdgcn_readlane(int v, int l) { return lx_emu_xchg(v, l); }
template <typename T> inline T __shfl_xor(T v, int mask) { return lx_emu_xchg(v, (threadIdx.x & 63) ^ mask); }

// reciprocal square root estimates: the FP64 one at single precision only (a
// coarser seed than the hardware's), so the emulated kernels show that their
// Newton steps reach full precision from it
inline double __builtin_amdgcn_rsq(double x) { return (double)(float)(1.0 / std::sqrt(x)); }
inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
inline void __builtin_amdgcn_s_sleep(int) { std::this_thread::yield(); }
// lanes are separate threads here, so a wave barrier must really synchronise them
inline void __builtin_amdgcn_wave_barrier() { lx_emu_wave_barrier[threadIdx.x >> 6].wait(); }
inline void __builtin_amdgcn_sched_barrier(int) {}
inline void __builtin_amdgcn_fence_emu() { std::atomic_thread_fence(std::memory_order_seq_cst); }
#define __builtin_amdgcn_fence(order, scope) __builtin_amdgcn_fence_emu()
inline void __builtin_amdgcn_s_waitcnt(int) {}
inline long long wall_clock64()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
               .count() / 10;
}
// __hip_atomic_load / _store / _fetch_add are clang builtins on every target;
// their scope argument takes clang's __MEMORY_SCOPE_* values.
#define __HIP_MEMORY_SCOPE_AGENT __MEMORY_SCOPE_DEVICE
#define __HIP_MEMORY_SCOPE_WORKGROUP __MEMORY_SCOPE_WRKGRP
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned *p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int *p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline int atomicMin(int *p, int v)
{
    int cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
using std::sqrt;
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
