#include <type_traits>
// CPU execution of the library's persistent executor (test infrastructure).
// Builds a plan with the product planner and runs the product kernel source
// (lx_exec_kernel.hpp) as one workgroup of 256 host threads, through the
// emulation header in this directory.  Exposes a C entry point for ctypes.
#include "hip/hip_runtime.h"

#include <climits>
#include <thread>
#include <vector>

#include "lx_exec_kernel.hpp"
#include "lx_planner.hpp"

// the sweep kernels' device code (the .hip file also holds host code that
// needs the HIP runtime; compile only its kernels here)
#define LX_SWEEP_KERNELS_ONLY
#include "lx_sweep.hip"
#define LX_BATCH_KERNELS_ONLY
#include "lx_batch.hip"
#define LX_ADJOINT_KERNELS_ONLY
#include "lx_adjoint.hip"

thread_local lx_emu_dim3 threadIdx, blockIdx;
lx_emu_dim3 blockDim, gridDim;
lx_emu_barrier lx_emu_wg_barrier;
lx_emu_barrier lx_emu_wave_barrier[4];
lx_emu_wave_buf lx_emu_wbuf[4];
lx_emu_bf_buf lx_emu_bfbuf[4];
static int g_update_prec = 0;
static int g_planes = 0;
static int g_wide_m = 128, g_wide_n = 512, g_trsm_wide = 64, g_workers = 512;

// unit shapes of the next lx_emu_run plans: GEMM units up to wide_m x wide_n
// off the lookahead column, panel solves of trsm_wide rows off the lookahead
// tile, the workgroup count the unit heights are chosen for
extern "C" void lx_emu_set_units(int wide_m, int wide_n, int trsm_wide, int workers)
{
    g_wide_m = wide_m;
    g_wide_n = wide_n;
    g_trsm_wide = trsm_wide;
    g_workers = workers;
}

// trailing-update precision of the next lx_emu_run plans (0 FP32, 1 split BF16)
extern "C" void lx_emu_set_update_prec(int u) { g_update_prec = u; }
static int g_group = 1;  // RL plans: panels per trailing-update pass
extern "C" void lx_emu_set_update_group(int v) { g_group = v; }
// 1: FP32 factors with split-FP16 updates get the plane buffer (TRSM units
// write the FP16 planes, eligible GEMM / SYRK units read them), as the
// library sets it up
extern "C" void lx_emu_set_planes(int on) { g_planes = on; }

extern "C" int lx_emu_run(int dtype, int what, int n, int b, int nrhs, void *A, long long lda, void *X, long long ldx,
                          int policy, int *info_out, int *done_out, int *nunits_out)
{
    lx_plan_params P;
    P.unit_m = LX_UNIT_M;
    P.unit_n = LX_UNIT_N;
    P.policy = policy;
    P.update_prec = dtype == 1 ? g_update_prec : 0;
    P.gemm_wide_m = g_wide_m;
    P.gemm_wide_n = g_wide_n;
    P.trsm_rows_wide = g_trsm_wide;
    P.workers = g_workers;
    P.update_group = g_group;
    lx_planner pl(P);
    int s = pl.add_system(n, b, nrhs);
    switch (what) {
    case 0: pl.factor_rl(s); break;
    case 1: pl.factor_ll(s); break;
    case 2: pl.factor_rl(s); pl.sweep_forward(s); pl.sweep_backward(s); break;
    case 3: pl.sweep_forward(s); pl.sweep_backward(s); break;
    case 4: pl.sweep_forward(s); break;
    case 5: pl.sweep_backward(s); break;
    default: return -2;
    }
    pl.finalize();
    const std::vector<lx_unit> &U = pl.units();
    std::vector<int> ctr(pl.num_counters(), 0);
    lx_status st = {};
    int info = INT_MAX;
    lx_sys sys = {};
    sys.A = A;
    sys.X = X;
    sys.lda = lda;
    sys.ldx = ldx;
    sys.n = n;
    sys.nrhs = nrhs;
    sys.ldw = (int64_t)n + 64;
    std::vector<double> wbuf((size_t)sys.ldw * 64);  // large enough for either precision
    sys.W = wbuf.data();
    std::vector<_Float16> planes;
    if (g_planes && dtype == 1 && P.update_prec == 2 && b % 16 == 0 && what <= 1) {
        sys.ldp = ((int64_t)n + 256 + 31) / 32 * 32;
        sys.nsig = (n + 15) / 16;
        planes.assign((size_t)2 * sys.nsig * sys.ldp * 16, (_Float16)0.0f);
        sys.planes = planes.data();
    }
    blockDim.x = 256;
    gridDim.x = 1;
    lx_emu_wg_barrier.init(256, "workgroup");
    for (int w = 0; w < 4; ++w) lx_emu_wave_barrier[w].init(64, "wavefront");
    const long long ticks = 100000000LL * 20;  // 20 s at the emulated 100 MHz clock
    std::vector<std::thread> th;
    th.reserve(256);
    for (int t = 0; t < 256; ++t)
        th.emplace_back([&, t] {
            threadIdx.x = t;
            blockIdx.x = 0;
            if (dtype == 0)
                lx_exec_kernel<double, LX_FAM_ALL>(U.data(), (int)U.size(), &sys, ctr.data(), &st, &info, ticks, nullptr);
            else
                lx_exec_kernel<float, LX_FAM_ALL>(U.data(), (int)U.size(), &sys, ctr.data(), &st, &info, ticks, nullptr);
        });
    for (auto &t : th) t.join();
    *info_out = info == INT_MAX ? 0 : info;
    *done_out = st.done;
    *nunits_out = (int)U.size();
    return (st.abort || st.timeouts) ? -1 : 0;
}

// Whether the factor plan of an n x n matrix (n a multiple of b) cuts every
// unit inside a b x b tile, with the library's tuning parameters given.
extern "C" int lx_emu_plan_in_tiles(int variant, int n, int b, int update_prec, int potrf_sub, int wide_n)
{
    lx_plan_params P;
    P.unit_m = LX_UNIT_M;
    P.unit_n = LX_UNIT_N;
    P.update_prec = update_prec;
    P.potrf_sub = potrf_sub;
    P.gemm_wide_n = wide_n;
    lx_planner pl(P);
    int s = pl.add_system(n, b, 0);
    if (variant == 0) pl.factor_rl(s);
    else pl.factor_ll(s);
    pl.finalize();
    return lx_units_in_tiles(pl.units(), b) ? 1 : 0;
}

// The factor of a hypermatrix (tile (i, j) at tiles[j * mt + i], b x b, ld b)
// through the tiled executor instantiation, one workgroup of 256 host threads.
extern "C" int lx_emu_run_tiles(int dtype, int variant, int mt, int b, void **tiles, int *info_out)
{
    const int n = mt * b;
    lx_plan_params P;
    P.unit_m = LX_UNIT_M;
    P.unit_n = LX_UNIT_N;
    P.update_prec = dtype == 1 ? g_update_prec : 0;
    lx_planner pl(P);
    int s = pl.add_system(n, b, 0);
    if (variant == 0) pl.factor_rl(s);
    else pl.factor_ll(s);
    pl.finalize();
    const std::vector<lx_unit> &U = pl.units();
    if (!lx_units_in_tiles(U, b)) return -3;
    std::vector<int> ctr(pl.num_counters(), 0);
    lx_status st = {};
    int info = INT_MAX;
    lx_sys sys = {};
    sys.lda = n;
    sys.n = n;
    sys.tiles = tiles;
    sys.tmt = mt;
    sys.tb = b;
    sys.ldw = (int64_t)n + 64;
    std::vector<double> wbuf((size_t)sys.ldw * 64);
    sys.W = wbuf.data();
    blockDim.x = 256;
    gridDim.x = 1;
    lx_emu_wg_barrier.init(256, "workgroup");
    for (int w = 0; w < 4; ++w) lx_emu_wave_barrier[w].init(64, "wavefront");
    const long long ticks = 100000000LL * 20;
    std::vector<std::thread> th;
    th.reserve(256);
    for (int t = 0; t < 256; ++t)
        th.emplace_back([&, t] {
            threadIdx.x = t;
            blockIdx.x = 0;
            if (dtype == 0)
                lx_exec_kernel<double, LX_FAM_FACTOR, true>(U.data(), (int)U.size(), &sys, ctr.data(), &st, &info,
                                                            ticks, nullptr);
            else
                lx_exec_kernel<float, LX_FAM_FACTOR, true>(U.data(), (int)U.size(), &sys, ctr.data(), &st, &info,
                                                           ticks, nullptr);
        });
    for (auto &t : th) t.join();
    *info_out = info == INT_MAX ? 0 : info;
    return (st.abort || st.timeouts || st.done != (int)U.size()) ? -1 : 0;
}

template <typename F>
static void run_wg(F &&body)
{
    lx_emu_wg_barrier.init(256, "workgroup");
    for (int w = 0; w < 4; ++w) lx_emu_wave_barrier[w].init(64, "wavefront");
    std::vector<std::thread> th;
    th.reserve(256);
    for (int t = 0; t < 256; ++t)
        th.emplace_back([&, t] {
            threadIdx.x = t;
            blockIdx.x = 0;
            body();
        });
    for (auto &t : th) t.join();
}

// The RL factor plan run with the first dependency counter of unit `stuck`
// (a unit with dependencies) started far below any threshold, so that its
// wait can only end through the bounded-wait path: spin_ticks later the wave
// sets the abort flag and every wave leaves the persistent loop.  Returns the
// unit's index, or -1 when no unit from `stuck` on has a dependency;
// st_out = {abort, timeouts, done, nunits}.
extern "C" int lx_emu_run_stuck(int dtype, int n, int b, int stuck, long long spin_ticks, void *A, long long lda,
                                int *st_out)
{
    lx_plan_params P;
    P.unit_m = LX_UNIT_M;
    P.unit_n = LX_UNIT_N;
    lx_planner pl(P);
    int s = pl.add_system(n, b, 0);
    pl.factor_rl(s);
    pl.finalize();
    const std::vector<lx_unit> &U = pl.units();
    int u = stuck;
    while (u < (int)U.size() && U[u].nd == 0) ++u;
    if (u >= (int)U.size()) return -1;
    std::vector<int> ctr(pl.num_counters(), 0);
    ctr[U[u].dc[0]] = -1000000000;
    lx_status st = {};
    int info = INT_MAX;
    lx_sys sys = {};
    sys.A = A;
    sys.lda = lda;
    sys.n = n;
    sys.ldw = (int64_t)n + 64;
    std::vector<double> wbuf((size_t)sys.ldw * 64);
    sys.W = wbuf.data();
    blockDim.x = 256;
    gridDim.x = 1;
    run_wg([&] {
        if (dtype == 0)
            lx_exec_kernel<double, LX_FAM_FACTOR>(U.data(), (int)U.size(), &sys, ctr.data(), &st, &info, spin_ticks,
                                                  nullptr);
        else
            lx_exec_kernel<float, LX_FAM_FACTOR>(U.data(), (int)U.size(), &sys, ctr.data(), &st, &info, spin_ticks,
                                                 nullptr);
    });
    st_out[0] = st.abort;
    st_out[1] = st.timeouts;
    st_out[2] = st.done;
    st_out[3] = (int)U.size();
    return u;
}

// C = alpha op(A) op(B) + beta C through lx_gemm64_kernel<float> (no split-K),
// uplo 0 full, 1 lower, 2 upper.
extern "C" void lx_emu_gemm64s(int ta, int tb, int M, int N, int K, float alpha, const float *A, long long lda,
                               const float *B, long long ldb, float beta, float *C, long long ldc, int uplo)
{
    const int tm = (M + 63) / 64, tn = (N + 63) / 64, kchunk = (K + 15) / 16 * 16;
    gridDim.x = tm;
    gridDim.y = tn;
    gridDim.z = 1;
    blockDim.x = 256;
    for (int y = 0; y < tn; ++y)
        for (int x = 0; x < tm; ++x)
            run_wg([&] {
                blockIdx.x = x;
                blockIdx.y = y;
                blockIdx.z = 0;
                if (ta == 0 && tb == 0)
                    lx_gemm64_kernel<float, 0, 0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, uplo, 1, kchunk, nullptr);
                else if (ta == 0)
                    lx_gemm64_kernel<float, 0, 1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, uplo, 1, kchunk, nullptr);
                else if (tb == 0)
                    lx_gemm64_kernel<float, 1, 0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, uplo, 1, kchunk, nullptr);
                else
                    lx_gemm64_kernel<float, 1, 1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, uplo, 1, kchunk, nullptr);
            });
}

// One FP32 GEMM / SYRK unit through the executor's dispatch (lx_run_unit):
// C (M x N, ldc) -= A B^T, A = rows 0..M-1 and B = rows M..M+N-1 of the K
// columns of P (ldp); flags = the unit's aux0 (bit 0 lower mask, bit 1
// split-BF16 bodies, bit 2 128 x 256 split-BF16 blocks), gdiag = its aux1.
extern "C" void lx_emu_gemm_unit(int flags, int M, int N, int K, float *C, long long ldc, const float *P,
                                 long long ldp, int gdiag)
{
    lx_sys sys = {};
    sys.A = C;
    sys.lda = ldc;
    sys.X = const_cast<float *>(P);
    sys.ldx = ldp;
    lx_unit u = {};
    u.op = (flags & 1) ? LX_OP_SYRK : LX_OP_GEMM_NT;
    u.M = M;
    u.N = N;
    u.K = K;
    u.aux0 = flags;
    u.aux1 = gdiag;
    u.cbase = LX_BASE_A;
    u.abase = u.bbase = LX_BASE_X;
    u.c = 0;
    u.a = 0;
    u.b = M;  // row M, column 0
    int info = INT_MAX;
    std::vector<double> smem((lx_fam_smem<float, LX_FAM_FACTOR>::bytes + 7) / 8);
    blockDim.x = 256;
    gridDim.x = 1;
    run_wg([&] { lx_run_unit<float, LX_FAM_FACTOR>(u, &sys, &info, reinterpret_cast<char *>(smem.data())); });
}

// The same unit with the plane buffer set: the operand rows P (rows 0..M-1 =
// A, M..M+N-1 = B, K columns, K a multiple of 16) are given only as FP16
// planes (hi, lo of 2^12 x, the split the TRSM units write), while the FP32
// operands the unit could also read are zeros: C changes only if the unit
// took its operands from the planes.
extern "C" void lx_emu_gemm_unit_planes(int flags, int M, int N, int K, float *C, long long ldc, const float *P,
                                        long long ldpm, int gdiag)
{
    lx_sys sys = {};
    std::vector<float> zeros((size_t)(M + N) * K, 0.0f);
    sys.A = C;
    sys.lda = ldc;
    sys.X = zeros.data();
    sys.ldx = M + N;
    sys.ldp = ((int64_t)(M + N) + 256 + 31) / 32 * 32;
    sys.nsig = K / 16;
    std::vector<_Float16> planes((size_t)2 * sys.nsig * sys.ldp * 16, (_Float16)0.0f);
    for (int r = 0; r < M + N; ++r)
        for (int c = 0; c < K; ++c) {
            const float xs = P[(size_t)c * ldpm + r] * 4096.0f;
            const _Float16 h = (_Float16)xs;
            const size_t o = ((size_t)(c / 16) * sys.ldp + r) * 16 + c % 16;
            planes[o] = h;
            planes[o + (size_t)sys.nsig * sys.ldp * 16] = (_Float16)(xs - (float)h);
        }
    sys.planes = planes.data();
    lx_unit u = {};
    u.op = (flags & 1) ? LX_OP_SYRK : LX_OP_GEMM_NT;
    u.M = M;
    u.N = N;
    u.K = K;
    u.aux0 = flags;
    u.aux1 = gdiag;
    u.cbase = LX_BASE_A;
    u.abase = u.bbase = LX_BASE_X;
    u.c = 0;
    u.a = 0;
    u.b = M;  // row M, column 0
    int info = INT_MAX;
    std::vector<double> smem((lx_fam_smem<float, LX_FAM_FACTOR>::bytes + 7) / 8);
    blockDim.x = 256;
    gridDim.x = 1;
    run_wg([&] { lx_run_unit<float, LX_FAM_FACTOR>(u, &sys, &info, reinterpret_cast<char *>(smem.data())); });
}

// One sweep (trans 0: L Y = X, 1: L^T Z = X) through lx_sweep_kernel, one
// workgroup at a time in ticket order.
// One sweep through lx_sweep_kernel.  stuck >= 0 starts the partial-sum
// counter of that step far below its count, so the step's solve unit can only
// leave its wait through the bounded-wait path (spin_ticks); returns -3 when
// a wait gave up (the abort flag), -1 / -2 when the counters or the ticket do
// not add up.
extern "C" int lx_emu_sweep_ex(int dtype, int trans, int n, int nrhs, const void *L, long long ldl, void *X,
                               long long ldx, int stuck, long long spin_ticks)
{
    // one emulated workgroup takes every unit in list order
    const int nb = (n + 63) / 64, C = 3;  // small chunks: several partial-sum units per step
    std::vector<int> units, rowbase;
    lx_sw_build_units(nb, C, units, rowbase);
    std::vector<int> ctl(64 + 2 * nb, 0);
    if (stuck >= 0 && stuck < nb) ctl[64 + nb + stuck] = -1000000000;
    std::vector<double> part((size_t)std::max(1, rowbase[nb]) * LX_SW_NR * 64);  // large enough for either type
    const int64_t ldmb = (int64_t)nb * 64;
    std::vector<unsigned long long> mbv((size_t)ldmb * LX_SW_NR, ~0ull);  // all-ones: empty (either type)
    const long long ticks = spin_ticks;
    auto go = [&](auto tp, auto tr, auto nr) {
        using T = decltype(tp);
        run_wg([&] {
            lx_sweep_kernel<T, decltype(tr)::value, decltype(nr)::value>(
                n, nrhs, (const T *)L, ldl, (T *)X, ldx, units.data(), (int)units.size(), rowbase.data(), C,
                (T *)part.data(), ctl.data(), (T *)mbv.data(), ldmb, ticks);
        });
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    auto by_nr = [&](auto tp, auto tr) {
        if (nrhs <= 1) go(tp, tr, std::integral_constant<int, 1>());
        else if (nrhs <= 4) go(tp, tr, std::integral_constant<int, 4>());
        else go(tp, tr, std::integral_constant<int, 16>());
    };
    if (dtype == 0) { if (trans) by_nr(double(), I1()); else by_nr(double(), I0()); }
    else { if (trans) by_nr(float(), I1()); else by_nr(float(), I0()); }
    if (ctl[1] != 0) return -3;  // a wait timed out
    for (int b = 0; b < nb; ++b)
        if (ctl[64 + nb + b] != lx_sw_nparts(b, C)) return -1;
    return ctl[0] == (int)units.size() + 1 ? 0 : -2;
}

extern "C" int lx_emu_sweep(int dtype, int trans, int n, int nrhs, const void *L, long long ldl, void *X,
                            long long ldx)
{
    return lx_emu_sweep_ex(dtype, trans, n, nrhs, L, ldl, X, ldx, -1, 100000000LL * 20);
}

// One FP32 sweep with 128-row steps: the step matrices through
// lx_sw2_prep_kernel (grid nb, each direction, one workgroup at a time), then
// lx_sweep2_kernel as one workgroup taking every unit in list order.
// Returns 0, -3 when a wait gave up, -1 when a partial sum was never written,
// -2 when the ticket does not add up.
extern "C" int lx_emu_sweep2(int trans, int n, int nrhs, const float *L, long long ldl, float *X, long long ldx,
                             int C)
{
    const int nb = (n + LX_S2 - 1) / LX_S2, W = nrhs > 1 ? 16 : 1;
    std::vector<float> prep((size_t)6 * LX_S2_MAT * nb);
    gridDim.x = nb;
    gridDim.y = 1;
    gridDim.z = 1;
    blockDim.x = 256;
    for (int y = 0; y < 2; ++y)
        for (int x = 0; x < nb; ++x)
            run_wg([&] {
                blockIdx.x = x;
                blockIdx.y = 0;
                blockIdx.z = 0;
                if (y) lx_sw2_prep_kernel<float, 1>(n, L, ldl, prep.data());
                else lx_sw2_prep_kernel<float, 0>(n, L, ldl, prep.data());
            });
    gridDim.x = 1;
    std::vector<int> units, rowbase;
    lx_s2_build_units(nb, C, units, rowbase);
    std::vector<int> ctl(64, 0);
    std::vector<unsigned int> part((size_t)std::max(1, rowbase[nb]) * LX_S2 * W, ~0u);  // all-ones: empty
    const int64_t ldmb = (int64_t)nb * LX_S2;
    std::vector<unsigned int> mbv((size_t)ldmb * W, ~0u);
    const long long ticks = 100000000LL * 20;
    run_wg([&] {
        if (nrhs > 1) {
            if (trans)
                lx_sweep2m_kernel<float, 1>(n, nrhs, L, ldl, X, ldx, prep.data(), units.data(), (int)units.size(),
                                            rowbase.data(), C, (float *)part.data(), ctl.data(), (float *)mbv.data(),
                                            ldmb, ticks);
            else
                lx_sweep2m_kernel<float, 0>(n, nrhs, L, ldl, X, ldx, prep.data(), units.data(), (int)units.size(),
                                            rowbase.data(), C, (float *)part.data(), ctl.data(), (float *)mbv.data(),
                                            ldmb, ticks);
        } else if (trans)
            lx_sweep2_kernel<float, 1, 1>(n, L, ldl, X, prep.data(), units.data(), (int)units.size(), rowbase.data(),
                                          C, (float *)part.data(), ctl.data(), (float *)mbv.data(), ticks);
        else
            lx_sweep2_kernel<float, 0, 1>(n, L, ldl, X, prep.data(), units.data(), (int)units.size(), rowbase.data(),
                                          C, (float *)part.data(), ctl.data(), (float *)mbv.data(), ticks);
    });
    if (ctl[1] != 0) return -3;
    for (int p = 0; p < rowbase[nb]; ++p)  // every partial sum written, in its first nrhs columns
        for (int q = 0; q < LX_S2 * std::min(W, nrhs); ++q)
            if (part[(size_t)p * LX_S2 * W + q] == ~0u) return -1;
    return ctl[0] == (int)units.size() + 1 ? 0 : -2;
}

// Panel mask of the next lx_emu_batch runs (bit c: orders 64c+1 .. 64c+64
// factor in 128-column panels).
static unsigned g_batch_pw_mask = 0;
extern "C" void lx_emu_set_batch_panels(unsigned mask) { g_batch_pw_mask = mask; }

// Batched refinement solves through lx_batch_ir_kernel, one workgroup taking
// every system from the queue.  stat: count x {sweeps, fallback, info, pad, berr}.
extern "C" int lx_emu_batch(int count, const int *n, const double *const *A, const long long *lda,
                            const double *const *B, double *const *X, const int *order, void *stat, int max_sweeps,
                            double tol, double stall)
{
    std::vector<lx_bsys> sys(count);
    int nmax = 1;
    for (int i = 0; i < count; ++i) {
        sys[i] = lx_bsys{A[i], B[i], X[i], (int64_t)lda[i], n[i], 0};
        nmax = std::max(nmax, n[i]);
    }
    const int64_t sb = lx_b_slot_bytes(nmax);
    std::vector<double> ws((size_t)sb / 8 + 1, 0.0);
    unsigned ticket = 0;
    gridDim.x = 1;
    blockDim.x = 256;
    run_wg([&] {
        lx_batch_ir_kernel<2>(sys.data(), order, count, (char *)ws.data(), sb, nmax, &ticket, (lx_bstat *)stat,
                           max_sweeps, tol, stall, nullptr, g_batch_pw_mask);
    });
    return ticket == (unsigned)count + 1 ? 0 : -1;
}

// The adjoint's diagonal-block kernel (mode 0: inverse, mode 1: level-2 adjoint), one workgroup.
extern "C" void lx_emu_adj_diag(int mode, int jb, const double *L, long long ldl, double *Ab, long long lda,
                                double *Zb, double *Sb)
{
    gridDim.x = 1;
    blockDim.x = 256;
    run_wg([&] { lx_adj_diag_kernel(mode, jb, L, ldl, Ab, lda, Zb, Sb, 64, 1.0); });
}

// The adjoint's GEMM kernel over its whole grid, one workgroup at a time.
extern "C" void lx_emu_adj_gemm(int ta, int tb, int M, int N, int K, double alpha, const double *A, long long lda,
                                const double *B, long long ldb, double beta, double *C, long long ldc, int lower,
                                int ksplit, int kchunk, double *part)
{
    const int tm = (M + 63) / 64, tn = (N + 63) / 64;
    gridDim.x = tm;
    gridDim.y = tn;
    gridDim.z = ksplit;
    blockDim.x = 256;
    for (int z = 0; z < ksplit; ++z)
        for (int y = 0; y < tn; ++y)
            for (int x = 0; x < tm; ++x)
                run_wg([&] {
                    blockIdx.x = x;
                    blockIdx.y = y;
                    blockIdx.z = z;
                    if (ta == 0 && tb == 0)
                        lx_gemm64_kernel<double, 0, 0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower, ksplit, kchunk, part);
                    else if (ta == 0)
                        lx_gemm64_kernel<double, 0, 1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower, ksplit, kchunk, part);
                    else if (tb == 0)
                        lx_gemm64_kernel<double, 1, 0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower, ksplit, kchunk, part);
                    else
                        lx_gemm64_kernel<double, 1, 1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower, ksplit, kchunk, part);
                });
}

// Positive control for the sanitizer build: element i of an 8-element heap
// array (i = 8 reads past its end, which AddressSanitizer must report).
extern "C" int lx_emu_heap_probe(int i)
{
    int *p = new int[8]();
    const int v = p[i];
    delete[] p;
    return v;
}
