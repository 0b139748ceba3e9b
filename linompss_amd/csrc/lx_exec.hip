// lx_exec.hip -- persistent tile-DAG executor and the factor / solve entry
// points of liblinompss.so.
//
// The reference runs its tile tasks through the OmpSs runtime: the master
// thread spawns them (src/cholesky/dense/chol_rlmain.c:25-50), Nanos++ worker
// threads execute them in priority order once their in/inout tiles are ready,
// and the caller waits with `#pragma omp taskwait` (chol.c:72).  Here one
// kernel launch replaces all of that: every workgroup loops
//   grab the next unit (one atomic) -> wait until each dependency counter
//   reaches its threshold -> run the tile kernel -> publish (release fence,
//   counter + 1)
// until the unit list is exhausted.  Units are listed in a topological order
// chosen by the host planner, so every wait is on a unit already held by a
// running workgroup and the launch cannot deadlock, whatever the occupancy.
// A wait that exceeds the timeout sets an abort flag that every workgroup
// observes, so each wave reaches the kernel's exit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "linompss.h"
#include "lx_exec_kernel.hpp"
#include "lx_internal.hpp"
#include "lx_planner.hpp"

__global__ void lx_fill_i32_kernel(int32_t *p, int32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

void lx_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(lx_fill_i32_kernel, dim3(blocks), dim3(256), 0, s, p, v, n);
}

// ---------------------------------------------------------------------------
// Host side: device context, plan cache, scratch.

namespace {

struct lx_dplan {
    std::string key;
    lx_unit *d_units = nullptr;
    int nunits = 0, ntasks = 0, nctr = 0;
    double sim_s = 0.0;
    bool in_tiles = false;  // every unit's operands inside single b x b tiles
};


struct lx_dev_ctx {
    int dev = -1;
    int cus = 0;
    int wall_khz = 0;
    int occ[2][3] = {};  // resident WGs per CU: [FP64 / FP32][op family]
    std::list<lx_dplan> plans;  // most recent first
    void *scratch[16] = {};
    size_t scratch_sz[16] = {};
};

std::mutex g_mu;
std::vector<std::unique_ptr<lx_dev_ctx>> g_ctx;

thread_local linompss_exec_stats g_last_stats = {};

lx_dev_ctx *ctx_get()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if ((int)g_ctx.size() <= dev) g_ctx.resize(dev + 1);
    if (!g_ctx[dev]) {
        auto c = std::make_unique<lx_dev_ctx>();
        c->dev = dev;
        if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return nullptr;
        if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
            c->wall_khz <= 0)
            c->wall_khz = 100000;
        int nb[6] = {};
        hipError_t e = hipSuccess;
        e = e ? e : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[0], lx_exec_kernel<double, LX_FAM_FACTOR>, LX_WG_THREADS, 0);
        e = e ? e : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[1], lx_exec_kernel<double, LX_FAM_SOLVE>, LX_WG_THREADS, 0);
        e = e ? e : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[2], lx_exec_kernel<double, LX_FAM_ALL>, LX_WG_THREADS, 0);
        e = e ? e : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[3], lx_exec_kernel<float, LX_FAM_FACTOR>, LX_WG_THREADS, 0);
        e = e ? e : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[4], lx_exec_kernel<float, LX_FAM_SOLVE>, LX_WG_THREADS, 0);
        e = e ? e : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[5], lx_exec_kernel<float, LX_FAM_ALL>, LX_WG_THREADS, 0);
        if (e != hipSuccess) return nullptr;
        for (int q = 0; q < 6; ++q) c->occ[q / 3][q % 3] = std::max(1, nb[q]);
        g_ctx[dev] = std::move(c);
    }
    return g_ctx[dev].get();
}

void *scratch_get(lx_dev_ctx *c, int slot, size_t bytes)
{
    if (slot < 0 || slot >= 16) return nullptr;
    if (c->scratch_sz[slot] >= bytes && c->scratch[slot]) return c->scratch[slot];
    if (c->scratch[slot]) {
        (void)hipDeviceSynchronize();
        (void)hipFree(c->scratch[slot]);
        c->scratch[slot] = nullptr;
        c->scratch_sz[slot] = 0;
    }
    size_t sz = std::max<size_t>(bytes, 256);
    if (hipMalloc(&c->scratch[slot], sz) != hipSuccess) {
        c->scratch[slot] = nullptr;
        return nullptr;
    }
    c->scratch_sz[slot] = sz;
    return c->scratch[slot];
}

// Factor configurations recorded by the autotuner (or the caller), per
// (dtype, variant, n, b, update_prec); guarded by g_mu.
struct lx_cfg_entry {
    int dtype, variant, n, b, update_prec;
    linompss_factor_config cfg;
};
std::vector<lx_cfg_entry> g_cfgs;

const lx_cfg_entry *cfg_find(int dtype, int variant, int n, int b, int update_prec)
{
    for (const auto &e : g_cfgs)
        if (e.dtype == dtype && e.variant == variant && e.n == n && e.b == b && e.update_prec == update_prec)
            return &e;
    return nullptr;
}

// Panels per trailing-update pass for tiles right of a group of steps:
// about 2048 columns (groups of 4 at b = 512, 2 at b = 1024).
int default_update_group(int b) { return b > 0 ? std::max(1, std::min(8, 2048 / b)) : 1; }

void fill_params(lx_plan_params &P, int dtype, int workers, const linompss_exec_opts *o,
                 const lx_shape *sh = nullptr, int b_hint = 0)
{
    const int tb = sh ? sh->b : b_hint;
    P.unit_m = LX_UNIT_M;
    P.unit_n = LX_UNIT_N;
    P.elem_bytes = dtype == 0 ? 8 : 4;
    P.workers = workers;
    if (dtype == 1) {
        P.flop_rate *= 2.0;
        P.potrf_rate *= 1.5;
        P.trsm_rate *= 1.5;
    }
    if (o) P.policy = o->policy;
    if (o && dtype == 1 && (o->update_prec == 1 || o->update_prec == 2)) P.update_prec = o->update_prec;
    // FP64: 256-wide diagonal sub-tiles for tiles of 512 and more (fewer,
    // longer steps on the critical path; measured at n = 32768, b = 512: 208.2
    // ms against 209.0 with 128), 128-wide below, where a 256-wide unit's own
    // trailing updates dominate it (n = 8192: b = 256 12.25 ms against 13.50,
    // b = 384 13.5 against 15.1)
    if (dtype == 0) P.potrf_sub = tb > 0 && tb < 512 ? 128 : 256;
    // FP64: 128-column GEMM units in the second half of the panels (more units
    // while the trailing matrix shrinks; 229.0 against 232.1 ms at n = 32768,
    // b = 512; the FP32 factor measured no gain at any tail length)
    if (dtype == 0) P.wide_tail = -1;
    // grouped trailing updates (each tile right of a group of steps read and
    // written once per group), about 2048 columns of panels per pass.  At
    // n = 32768: b = 512, groups of 1 / 2 / 3 / 4: FP32 factor with split-FP16
    // updates 98.4 / 79.0 / 73.9 / 72.6 ms, FP64 225.7 / 218.1 / 215.4 /
    // 213.1 ms; b = 1024, groups of 1 / 2 / 4: FP32 78.1 / 70.2 / 72.8 ms
    P.update_group = sh ? default_update_group(sh->b) : 1;
    // GEMM units taller than 128 rows while the step's parallelism allows,
    // panel solves of 128 rows off the lookahead tile: each publication writes
    // the XCD's L2 back.  Measured at n = 32768, b = 512: FP32 factor 113.4 ms
    // with 128-row GEMM units, 102.1 with 256, 100.9 with 512 everywhere, 99.3
    // with up to 512 where each step keeps two units per workgroup, 98.7 with
    // 128-row panel-solve units besides; FP64 226.5 ms with up to 256 rows
    // against 228.9 with 128 (229.7 with 512, no gain from taller solves)
    P.gemm_wide_m = dtype == 1 ? 512 : 256;
    if (dtype == 1) P.trsm_rows_wide = 128;
    // a configuration recorded for this factor's shape (autotuner)
    if (sh && (sh->what == LX_W_FACTOR_RL || sh->what == LX_W_FACTOR_LL || sh->what == LX_W_POSV_RL)) {
        const int variant = sh->what == LX_W_FACTOR_LL ? 1 : 0;
        if (const lx_cfg_entry *e = cfg_find(sh->dtype, variant, sh->n, sh->b, P.update_prec)) {
            P.potrf_sub = e->cfg.potrf_sub;
            P.gemm_wide_m = e->cfg.gemm_wide_m;
            P.gemm_wide_n = e->cfg.gemm_wide_n;
            P.trsm_rows_wide = e->cfg.trsm_rows_wide;
            P.wide_tail = e->cfg.wide_tail;
            P.update_group = e->cfg.update_group;
        }
    }
    // tuning knobs: the sub-tile of diagonal-tile factorisations, the width
    // of GEMM units off the lookahead column, 128 x 256 split-BF16 blocks
    if (const char *ev = getenv("LINOMPSS_POTRF_SUB")) {
        const int v = atoi(ev);
        if (v >= 64 && v % 64 == 0) P.potrf_sub = v;
    }
    if (const char *ev = getenv("LINOMPSS_WIDE_N")) {
        const int v = atoi(ev);
        if (v >= 128 && v % 128 == 0) P.gemm_wide_n = v;
    }
    if (const char *ev = getenv("LINOMPSS_BX_WIDE")) P.bx_wide = atoi(ev) != 0;
    if (const char *ev = getenv("LINOMPSS_WIDE_TAIL")) {
        const int v = atoi(ev);
        if (v >= -1) P.wide_tail = v;
    }
    if (const char *ev = getenv("LINOMPSS_WIDE_M")) {
        const int v = atoi(ev);
        if (v >= 128 && v % 128 == 0) P.gemm_wide_m = v;
    }
    if (const char *ev = getenv("LINOMPSS_TRSM_ROWS")) {
        const int v = atoi(ev);
        if (v >= 64 && v % 64 == 0) P.trsm_rows_wide = v;
    }
    if (const char *ev = getenv("LINOMPSS_WIDE_UNITS")) {
        const double v = atof(ev);
        if (v >= 0.0) P.wide_units = v;
    }
    if (const char *ev = getenv("LINOMPSS_UPDATE_GROUP")) {
        const int v = atoi(ev);
        if (v >= 1 && v <= 8) P.update_group = v;
    }
    // a hypermatrix's tiles are separate buffers: no unit may span two panels
    if (sh && sh->tiled) P.update_group = 1;
}

bool build_plan(lx_planner &pl, const lx_shape &sh)
{
    int s = pl.add_system(sh.n, sh.b, sh.nrhs);
    switch (sh.what) {
    case LX_W_FACTOR_RL: pl.factor_rl(s); break;
    case LX_W_FACTOR_LL: pl.factor_ll(s); break;
    case LX_W_POSV_RL:
        pl.factor_rl(s);
        pl.sweep_forward(s);
        pl.sweep_backward(s);
        break;
    case LX_W_SWEEPS:
        pl.sweep_forward(s);
        pl.sweep_backward(s);
        break;
    case LX_W_FWD: pl.sweep_forward(s); break;
    case LX_W_BWD: pl.sweep_backward(s); break;
    default: return false;
    }
    pl.finalize();
    return true;
}

// Returns a cached device plan (or builds and uploads one).
lx_dplan *plan_get(lx_dev_ctx *c, const lx_shape &sh, int workers, const linompss_exec_opts *o)
{
    lx_plan_params P;
    fill_params(P, sh.dtype, workers, o, &sh);
    char key[256];
    snprintf(key, sizeof key, "w%d d%d n%d b%d r%d p%d um%d un%d W%d u%d ps%d wn%d wm%d wu%g tr%d bw%d wt%d ug%d",
             sh.what, sh.dtype, sh.n, sh.b, sh.nrhs, P.policy, P.unit_m, P.unit_n, P.workers, P.update_prec,
             P.potrf_sub, P.gemm_wide_n, P.gemm_wide_m, P.wide_units, P.trsm_rows_wide, P.bx_wide, P.wide_tail,
             P.update_group);
    for (auto it = c->plans.begin(); it != c->plans.end(); ++it)
        if (it->key == key) {
            c->plans.splice(c->plans.begin(), c->plans, it);
            return &c->plans.front();
        }
    lx_planner pl(P);
    try {
        if (!build_plan(pl, sh)) return nullptr;
    } catch (...) {
        return nullptr;
    }
    lx_dplan d;
    d.key = key;
    d.nunits = (int)pl.units().size();
    d.ntasks = (int)pl.tasks().size();
    d.nctr = pl.num_counters();
    d.sim_s = pl.simulated_time();
    d.in_tiles = (sh.what == LX_W_FACTOR_RL || sh.what == LX_W_FACTOR_LL) && sh.n % sh.b == 0 &&
                 lx_units_in_tiles(pl.units(), sh.b);
    if (d.nunits > 0) {
        if (hipMalloc(&d.d_units, sizeof(lx_unit) * d.nunits) != hipSuccess) return nullptr;
        if (hipMemcpy(d.d_units, pl.units().data(), sizeof(lx_unit) * d.nunits, hipMemcpyHostToDevice) !=
            hipSuccess) {
            (void)hipFree(d.d_units);
            return nullptr;
        }
    }
    c->plans.push_front(d);
    while (c->plans.size() > 8) {
        (void)hipDeviceSynchronize();
        if (c->plans.back().d_units) (void)hipFree(c->plans.back().d_units);
        c->plans.pop_back();
    }
    return &c->plans.front();
}

}  // namespace

std::mutex &lx_dev_mutex() { return g_mu; }

void *lx_scratch(int slot, size_t bytes)
{
    lx_dev_ctx *c = ctx_get();
    return c ? scratch_get(c, slot, bytes) : nullptr;
}

// Scratch slots: 0 = executor control (status, counters, info, sys table),
// 1 = factor workspace (inverse diagonal blocks), 2 = executor profile,
// 3 = sweep flags, 4..7 = refinement workspaces, 8 = residual partial sums,
// 9 = adjoint workspace, 10 = log det, 11 = FP16 operand planes, 12 = GP
// gradient partials, 13 = step matrices of a 128-row sweep, 14 = the
// refinement loop's step matrices (kept for all its solves).
static bool lx_trace_on()
{
    static int on = -1;
    if (on < 0) on = getenv("LINOMPSS_TRACE") ? 1 : 0;
    return on == 1;
}
#define LX_TRACE(...)                                                                                               \
    do {                                                                                                            \
        if (lx_trace_on()) {                                                                                        \
            fprintf(stderr, "[linompss] " __VA_ARGS__);                                                             \
            fputc('\n', stderr);                                                                                    \
            fflush(stderr);                                                                                         \
        }                                                                                                           \
    } while (0)

int lx_run_single(const lx_shape &sh, void *dA, int64_t lda, void *dX, int64_t ldx, hipStream_t stream,
                  const linompss_exec_opts *opts, int *info_out, void *const *dtiles)
{
    LX_TRACE("run what=%d dtype=%d n=%d b=%d nrhs=%d", sh.what, sh.dtype, sh.n, sh.b, sh.nrhs);
    lx_dev_ctx *c = ctx_get();
    if (!c) return LINOMPSS_ERR_NODEV;
    const int fam = (sh.what == LX_W_FACTOR_RL || sh.what == LX_W_FACTOR_LL) ? LX_FAM_FACTOR
                    : sh.what == LX_W_POSV_RL                                 ? LX_FAM_ALL
                                                                              : LX_FAM_SOLVE;
    const int occ = c->occ[sh.dtype == 0 ? 0 : 1][fam];
    LX_TRACE("ctx %p cus=%d wall_khz=%d family=%d occ=%d", (void *)c, c->cus, c->wall_khz, fam, occ);
    int wpc = occ;
    if (opts && opts->wg_per_cu > 0) wpc = std::min(opts->wg_per_cu, occ);
    const int workers = c->cus * wpc;
    lx_shape shp = sh;
    shp.tiled = dtiles != nullptr;
    lx_dplan *pl = plan_get(c, shp, workers, opts);
    if (!pl) return LINOMPSS_ERR_PLAN;
    LX_TRACE("plan units=%d tasks=%d ctr=%d", pl->nunits, pl->ntasks, pl->nctr);
    if (dtiles && !pl->in_tiles) return LINOMPSS_ERR_PLAN;
    if (pl->nunits == 0) {
        *info_out = 0;
        return 0;
    }
    // control block: status | info | sys | counters
    const size_t off_info = 64, off_sys = 128, off_ctr = 256;
    const size_t bytes = off_ctr + sizeof(int32_t) * (size_t)pl->nctr;
    char *ctl = (char *)scratch_get(c, 0, bytes);
    if (!ctl) return LINOMPSS_ERR_NOMEM;
    lx_sys hs = {};
    hs.A = dA;
    hs.X = dX;
    hs.lda = lda;
    hs.ldx = ldx;
    hs.n = sh.n;
    hs.nrhs = sh.nrhs;
    hs.info_slot = 0;
    if (dtiles) {
        hs.tiles = dtiles;
        hs.tmt = sh.n / sh.b;
        hs.tb = sh.b;
    }
    // split-FP16 trailing updates of an FP32 factor: the operands' FP16 planes,
    // written by the TRSM units and read by the GEMM / SYRK units (slot 11)
    // (LINOMPSS_PLANES=0, read at each call, keeps the split in the GEMM units)
    const char *pe = getenv("LINOMPSS_PLANES");
    if (sh.dtype == 1 && fam == LX_FAM_FACTOR && opts && opts->update_prec == 2 && sh.b % 16 == 0 &&
        !(pe && atoi(pe) == 0)) {
        hs.ldp = ((int64_t)sh.n + 256 + 31) / 32 * 32;
        hs.nsig = (sh.n + 15) / 16;
        hs.planes = scratch_get(c, 11, (size_t)2 * hs.nsig * hs.ldp * 32);
        if (!hs.planes) return LINOMPSS_ERR_NOMEM;
    }
    if (fam != LX_FAM_SOLVE) {
        hs.ldw = (int64_t)sh.n + 64;
        hs.W = scratch_get(c, 1, (size_t)hs.ldw * 64 * (sh.dtype == 0 ? sizeof(double) : sizeof(float)));
        if (!hs.W) return LINOMPSS_ERR_NOMEM;
    }
    LX_HIP_TRY(hipMemcpyAsync(ctl + off_sys, &hs, sizeof hs, hipMemcpyHostToDevice, stream));
    lx_fill_i32((int32_t *)ctl, 0, 16, stream);
    lx_fill_i32((int32_t *)(ctl + off_info), INT_MAX, 1, stream);
    lx_fill_i32((int32_t *)(ctl + off_ctr), 0, pl->nctr, stream);
    double tmo = 30.0;
    if (const char *ev = getenv("LINOMPSS_TIMEOUT_S")) tmo = atof(ev) > 0 ? atof(ev) : tmo;
    if (opts && opts->timeout_s > 0) tmo = opts->timeout_s;
    const long long ticks = (long long)(tmo * c->wall_khz * 1000.0);
    const int grid = std::max(1, std::min(pl->nunits, workers));
    hipEvent_t e0, e1;
    LX_HIP_TRY(hipEventCreate(&e0));
    LX_HIP_TRY(hipEventCreate(&e1));
    LX_HIP_TRY(hipEventRecord(e0, stream));
    long long *prof = nullptr;
    static const bool prof_on = getenv("LINOMPSS_EXEC_PROFILE") != nullptr;
    if (prof_on) prof = (long long *)scratch_get(c, 2, sizeof(long long) * LX_PROF_WORDS * (size_t)grid);
#define LX_LAUNCH(TT, FF, ...)                                                                                     \
    hipLaunchKernelGGL((lx_exec_kernel<TT, FF, ##__VA_ARGS__>), dim3(grid), dim3(LX_WG_THREADS), 0, stream,           \
                       pl->d_units, pl->nunits, (const lx_sys *)(ctl + off_sys), (int *)(ctl + off_ctr),             \
                       (lx_status *)ctl, (int *)(ctl + off_info), ticks, prof)
    if (dtiles) {  // factor plans only (in_tiles is false otherwise)
        if (sh.dtype == 0) LX_LAUNCH(double, LX_FAM_FACTOR, true);
        else LX_LAUNCH(float, LX_FAM_FACTOR, true);
    } else if (sh.dtype == 0) {
        if (fam == LX_FAM_FACTOR) LX_LAUNCH(double, LX_FAM_FACTOR);
        else if (fam == LX_FAM_SOLVE) LX_LAUNCH(double, LX_FAM_SOLVE);
        else LX_LAUNCH(double, LX_FAM_ALL);
    } else {
        if (fam == LX_FAM_FACTOR) LX_LAUNCH(float, LX_FAM_FACTOR);
        else if (fam == LX_FAM_SOLVE) LX_LAUNCH(float, LX_FAM_SOLVE);
        else LX_LAUNCH(float, LX_FAM_ALL);
    }
#undef LX_LAUNCH
    LX_HIP_TRY(hipGetLastError());
    LX_TRACE("launched grid=%d", grid);
    LX_HIP_TRY(hipEventRecord(e1, stream));
    lx_status hst;
    int hinfo = 0;
    LX_HIP_TRY(hipMemcpyAsync(&hst, ctl, sizeof hst, hipMemcpyDeviceToHost, stream));
    LX_HIP_TRY(hipMemcpyAsync(&hinfo, ctl + off_info, sizeof hinfo, hipMemcpyDeviceToHost, stream));
    LX_HIP_TRY(hipStreamSynchronize(stream));
    LX_TRACE("done head=%d abort=%d timeouts=%d done=%d info=%d", hst.head, hst.abort, hst.timeouts, hst.done, hinfo);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_last_stats.exec_ms = ms;
    g_last_stats.units = pl->nunits;
    g_last_stats.tasks = pl->ntasks;
    g_last_stats.grid = grid;
    g_last_stats.sim_ms = pl->sim_s * 1e3;
    if (prof) {
        std::vector<long long> h((size_t)LX_PROF_WORDS * grid);
        if (hipMemcpy(h.data(), prof, sizeof(long long) * h.size(), hipMemcpyDeviceToHost) == hipSuccess) {
            double tot[LX_PROF_WORDS] = {};
            for (int g2 = 0; g2 < grid; ++g2)
                for (int q = 0; q < LX_PROF_WORDS; ++q) tot[q] += (double)h[(size_t)g2 * LX_PROF_WORDS + q];
            const double tick_ms = 1.0 / c->wall_khz;  // per workgroup, summed
            static const char *nm[LX_OP_NOP] = {"potrf", "trsm", "syrk", "gemm", "trsv", "gemv"};
            fprintf(stderr, "[linompss profile] what=%d dtype=%d n=%d grid=%d launch=%.2f ms  wait=%.1f%%", sh.what,
                    sh.dtype, sh.n, grid, ms, 100.0 * tot[0] * tick_ms / (ms * grid));
            for (int q = 0; q < LX_OP_NOP; ++q)
                if (tot[1 + LX_OP_NOP + q] > 0)
                    fprintf(stderr, "  %s: n=%.0f %.1f%% avg=%.1f us (publish %.1f)", nm[q], tot[1 + LX_OP_NOP + q],
                            100.0 * tot[1 + q] * tick_ms / (ms * grid), tot[1 + q] * tick_ms * 1e3 / tot[1 + LX_OP_NOP + q],
                            tot[1 + 2 * LX_OP_NOP + q] * tick_ms * 1e3 / tot[1 + LX_OP_NOP + q]);
            fprintf(stderr, "\n");
        }
    }
    if (hst.abort || hst.timeouts || hst.done != pl->nunits) return LINOMPSS_ERR_SCHED;
    *info_out = hinfo == INT_MAX ? 0 : hinfo;
    return 0;
}

// ---------------------------------------------------------------------------
// C ABI.

static int check_args(int n, int b, int64_t lda)
{
    if (n < 0) return -1;
    if (b <= 0) return -2;
    if (lda < std::max(1, n)) return -5;
    return 0;
}

extern "C" int linompss_chol_dev(int dtype, int variant, int n, int b, void *dA, int64_t lda, void *stream,
                                 const linompss_exec_opts *opts)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (variant != 0 && variant != 1) return -2;
    if (n < 0) return -3;
    if (b <= 0) return -4;
    if (lda < std::max(1, n)) return -6;
    if (n == 0) return 0;
    std::lock_guard<std::mutex> g(g_mu);
    lx_shape sh = {variant == 0 ? LX_W_FACTOR_RL : LX_W_FACTOR_LL, dtype, n, b, 0};
    int info = 0;
    int rc = lx_run_single(sh, dA, lda, nullptr, 0, (hipStream_t)stream, opts, &info);
    return rc ? rc : info;
}

extern "C" int linompss_posv_dev(int dtype, int n, int nrhs, int b, void *dA, int64_t lda, void *dX, int64_t ldx,
                                 void *stream, const linompss_exec_opts *opts)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (b <= 0) return -4;
    if (lda < std::max(1, n)) return -6;
    if (ldx < std::max(1, n)) return -8;
    if (n == 0) return 0;
    std::lock_guard<std::mutex> g(g_mu);
    const bool fused = opts && opts->fused_solve;
    lx_shape sh = {(nrhs > 0 && fused) ? LX_W_POSV_RL : LX_W_FACTOR_RL, dtype, n, b, nrhs};
    int info = 0;
    int rc = lx_run_single(sh, dA, lda, dX, ldx, (hipStream_t)stream, opts, &info);
    if (rc) return rc;
    if (nrhs > 0 && !fused && info == 0) {
        rc = lx_sweeps_dev(dtype, n, nrhs, dA, lda, dX, ldx, 3, (hipStream_t)stream);
        if (rc) return rc;
        LX_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    }
    return info;
}

extern "C" int linompss_potrs_dev(int dtype, int n, int nrhs, int b, const void *dL, int64_t lda, void *dX,
                                  int64_t ldx, void *stream, const linompss_exec_opts *opts)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (b <= 0) return -4;
    if (lda < std::max(1, n)) return -6;
    if (ldx < std::max(1, n)) return -8;
    if (n == 0 || nrhs == 0) return 0;
    std::lock_guard<std::mutex> g(g_mu);
    if (opts && opts->fused_solve) {  // the sweeps as a task DAG on the persistent executor
        lx_shape sh = {LX_W_SWEEPS, dtype, n, b, nrhs};
        int info = 0;
        return lx_run_single(sh, const_cast<void *>(dL), lda, dX, ldx, (hipStream_t)stream, opts, &info);
    }
    int rc = lx_sweeps_dev(dtype, n, nrhs, dL, lda, dX, ldx, 3, (hipStream_t)stream);
    if (rc) return rc;
    LX_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

extern "C" int linompss_trsm_dev(int dtype, int trans, int n, int nrhs, const void *dL, int64_t ldl, void *dX,
                                 int64_t ldx, void *stream)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (trans != 0 && trans != 1) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldl < std::max(1, n)) return -6;
    if (ldx < std::max(1, n)) return -8;
    if (n == 0 || nrhs == 0) return 0;
    std::lock_guard<std::mutex> g(g_mu);
    int rc = lx_sweeps_dev(dtype, n, nrhs, dL, ldl, dX, ldx, trans ? 2 : 1, (hipStream_t)stream);
    if (rc) return rc;
    LX_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

extern "C" int linompss_last_exec_stats(linompss_exec_stats *st)
{
    if (!st) return -1;
    *st = g_last_stats;
    return 0;
}

// Host-memory wrappers: stage through device memory with the caller's lda.
template <typename T>
static int host_chol(int variant, int n, int b, T *A, int lda)
{
    int a = check_args(n, b, lda);
    if (a) return a;
    if (n == 0) return 0;
    const int dtype = sizeof(T) == 8 ? 0 : 1;
    T *dA = nullptr;
    if (hipMalloc(&dA, sizeof(T) * (size_t)n * n) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    int rc = 0;
    if (hipMemcpy2D(dA, sizeof(T) * n, A, sizeof(T) * (size_t)lda, sizeof(T) * n, n, hipMemcpyHostToDevice) !=
        hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc) rc = linompss_chol_dev(dtype, variant, n, b, dA, n, nullptr, nullptr);
    if (rc >= 0 &&
        hipMemcpy2D(A, sizeof(T) * (size_t)lda, dA, sizeof(T) * n, sizeof(T) * n, n, hipMemcpyDeviceToHost) !=
            hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    (void)hipFree(dA);
    return rc;
}

extern "C" int ompss_dchol_rl(int n, int b, int t, double *A, int lda)
{
    (void)t;
    return host_chol<double>(0, n, b, A, lda);
}
extern "C" int ompss_dchol_ll(int n, int b, int t, double *A, int lda)
{
    (void)t;
    return host_chol<double>(1, n, b, A, lda);
}
extern "C" int ompss_schol_rl(int n, int b, int t, float *A, int lda)
{
    (void)t;
    return host_chol<float>(0, n, b, A, lda);
}
extern "C" int ompss_schol_ll(int n, int b, int t, float *A, int lda)
{
    (void)t;
    return host_chol<float>(1, n, b, A, lda);
}

// Hypermatrix A: the same plans and unit bodies, each unit addressed inside
// its tile (lx_run_unit); the plan is checked to cut every unit inside a tile.
extern "C" int linompss_chol_tiles_dev(int dtype, int variant, int mt, int b, void *const *dTiles, void *stream,
                                       const linompss_exec_opts *opts)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (variant != 0 && variant != 1) return -2;
    if (mt < 0 || (int64_t)mt * std::max(b, 1) > INT_MAX) return -3;
    if (b <= 0) return -4;
    if (mt == 0) return 0;
    if (!dTiles) return -5;
    std::lock_guard<std::mutex> g(g_mu);
    lx_shape sh = {variant == 0 ? LX_W_FACTOR_RL : LX_W_FACTOR_LL, dtype, mt * b, b, 0};
    int info = 0;
    int rc = lx_run_single(sh, nullptr, (int64_t)mt * b, nullptr, 0, (hipStream_t)stream, opts, &info, dTiles);
    return rc ? rc : info;
}

// Host tiles: the lower tiles are staged into one device buffer, the table of
// their device addresses uploaded, the factor run, and the tiles copied back.
template <typename T>
static int host_chol_tiles(int variant, int mt, int b, T **Ah)
{
    if (mt < 0) return -1;
    if (b <= 0) return -2;
    if (mt == 0) return 0;
    if (!Ah) return -4;
    for (int j = 0; j < mt; ++j)
        for (int i = j; i < mt; ++i)
            if (!Ah[(size_t)j * mt + i]) return -4;
    const int dtype = sizeof(T) == 8 ? 0 : 1;
    const size_t tsz = (size_t)b * b, nt = (size_t)mt * (mt + 1) / 2;
    T *dbuf = nullptr;
    void **dtab = nullptr;
    if (hipMalloc(&dbuf, sizeof(T) * tsz * nt) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    if (hipMalloc(&dtab, sizeof(void *) * (size_t)mt * mt) != hipSuccess) {
        (void)hipFree(dbuf);
        return LINOMPSS_ERR_NOMEM;
    }
    std::vector<void *> tab((size_t)mt * mt, nullptr);
    int rc = 0;
    size_t q = 0;
    for (int j = 0; j < mt; ++j)
        for (int i = j; i < mt; ++i) {
            T *d = dbuf + tsz * q++;
            tab[(size_t)j * mt + i] = d;
            if (!rc && hipMemcpy(d, Ah[(size_t)j * mt + i], sizeof(T) * tsz, hipMemcpyHostToDevice) != hipSuccess)
                rc = LINOMPSS_ERR_HIP;
        }
    if (!rc && hipMemcpy(dtab, tab.data(), sizeof(void *) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc) rc = linompss_chol_tiles_dev(dtype, variant, mt, b, dtab, nullptr, nullptr);
    if (rc >= 0)
        for (int j = 0; j < mt && rc >= 0; ++j)
            for (int i = j; i < mt; ++i)
                if (hipMemcpy(Ah[(size_t)j * mt + i], tab[(size_t)j * mt + i], sizeof(T) * tsz,
                              hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = LINOMPSS_ERR_HIP;
                    break;
                }
    (void)hipFree(dtab);
    (void)hipFree(dbuf);
    return rc;
}

extern "C" int ompss_dchol_hrl(int mt, int b, int t, double **Ah)
{
    (void)t;
    return host_chol_tiles<double>(0, mt, b, Ah);
}
extern "C" int ompss_dchol_hll(int mt, int b, int t, double **Ah)
{
    (void)t;
    return host_chol_tiles<double>(1, mt, b, Ah);
}
extern "C" int ompss_schol_hrl(int mt, int b, int t, float **Ah)
{
    (void)t;
    return host_chol_tiles<float>(0, mt, b, Ah);
}
extern "C" int ompss_schol_hll(int mt, int b, int t, float **Ah)
{
    (void)t;
    return host_chol_tiles<float>(1, mt, b, Ah);
}
// The nested variants sub-block each tile task by t on the CPU; here the
// executor cuts its own units, so they run the same factor.
extern "C" int ompss_dchol_nhrl(int mt, int b, int t, double **Ah)
{
    (void)t;
    return host_chol_tiles<double>(0, mt, b, Ah);
}
extern "C" int ompss_dchol_nhll(int mt, int b, int t, double **Ah)
{
    (void)t;
    return host_chol_tiles<double>(1, mt, b, Ah);
}
extern "C" int ompss_schol_nhrl(int mt, int b, int t, float **Ah)
{
    (void)t;
    return host_chol_tiles<float>(0, mt, b, Ah);
}
extern "C" int ompss_schol_nhll(int mt, int b, int t, float **Ah)
{
    (void)t;
    return host_chol_tiles<float>(1, mt, b, Ah);
}

template <typename T>
static int host_spdsolv(int m, int n, int b, T *A, T *B)
{
    if (m < 0) return -1;
    if (n < 0) return -2;
    if (b <= 0) return -3;
    if (m == 0) return 0;
    const int dtype = sizeof(T) == 8 ? 0 : 1;
    T *dA = nullptr, *dB = nullptr;
    if (hipMalloc(&dA, sizeof(T) * (size_t)m * m) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    if (n > 0 && hipMalloc(&dB, sizeof(T) * (size_t)m * n) != hipSuccess) {
        (void)hipFree(dA);
        return LINOMPSS_ERR_NOMEM;
    }
    int rc = 0;
    if (hipMemcpy(dA, A, sizeof(T) * (size_t)m * m, hipMemcpyHostToDevice) != hipSuccess) rc = LINOMPSS_ERR_HIP;
    if (!rc && n > 0 && hipMemcpy(dB, B, sizeof(T) * (size_t)m * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc) rc = linompss_posv_dev(dtype, m, n, b, dA, m, dB ? (void *)dB : (void *)dA, m, nullptr, nullptr);
    if (rc >= 0) {
        if (hipMemcpy(A, dA, sizeof(T) * (size_t)m * m, hipMemcpyDeviceToHost) != hipSuccess) rc = LINOMPSS_ERR_HIP;
        if (n > 0 && hipMemcpy(B, dB, sizeof(T) * (size_t)m * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = LINOMPSS_ERR_HIP;
    }
    (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    return rc;
}

extern "C" int dspdsolv_main(int m, int n, int b, double *A, double *B) { return host_spdsolv<double>(m, n, b, A, B); }
extern "C" int sspdsolv_main(int m, int n, int b, float *A, float *B) { return host_spdsolv<float>(m, n, b, A, B); }

template <typename T>
__global__ void lx_scale_kernel(T *B, int64_t ldb, int m, int n, T alpha)
{
    int64_t tot = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / m, r = e % m;
        B[c * ldb + r] *= alpha;
    }
}

template <typename T>
static int host_ltrsm(char uplo, char trans, char diag, int bm, int bn, int m, int n, T alpha, const T *A, int lda,
                      T *B, int ldb)
{
    (void)n;
    if (uplo != 'L' && uplo != 'l') return -1;
    if (trans != 'N' && trans != 'n' && trans != 'T' && trans != 't') return -2;
    if (diag != 'N' && diag != 'n') return -3;
    if (bm <= 0) return -4;
    if (bn < 0) return -5;
    if (m < 0) return -6;
    if (lda < std::max(1, m)) return -10;
    if (ldb < std::max(1, m)) return -12;
    if (m == 0 || bn == 0) return 0;
    const int dtype = sizeof(T) == 8 ? 0 : 1;
    const bool tr = trans == 'T' || trans == 't';
    T *dA = nullptr, *dB = nullptr;
    if (hipMalloc(&dA, sizeof(T) * (size_t)m * m) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    if (hipMalloc(&dB, sizeof(T) * (size_t)m * bn) != hipSuccess) {
        (void)hipFree(dA);
        return LINOMPSS_ERR_NOMEM;
    }
    int rc = 0;
    if (hipMemcpy2D(dA, sizeof(T) * m, A, sizeof(T) * (size_t)lda, sizeof(T) * m, m, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy2D(dB, sizeof(T) * m, B, sizeof(T) * (size_t)ldb, sizeof(T) * m, bn, hipMemcpyHostToDevice) !=
            hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    if (!rc && alpha != T(1)) {
        hipLaunchKernelGGL(lx_scale_kernel<T>, dim3(256), dim3(256), 0, nullptr, dB, (int64_t)m, m, bn, alpha);
        if (hipGetLastError() != hipSuccess) rc = LINOMPSS_ERR_HIP;
    }
    if (!rc) {
        std::lock_guard<std::mutex> g(g_mu);
        rc = lx_sweeps_dev(dtype, m, bn, dA, m, dB, m, tr ? 2 : 1, nullptr);
        if (!rc && hipDeviceSynchronize() != hipSuccess) rc = LINOMPSS_ERR_HIP;
    }
    if (!rc &&
        hipMemcpy2D(B, sizeof(T) * (size_t)ldb, dB, sizeof(T) * m, sizeof(T) * m, bn, hipMemcpyDeviceToHost) !=
            hipSuccess)
        rc = LINOMPSS_ERR_HIP;
    (void)hipFree(dA);
    (void)hipFree(dB);
    return rc;
}

extern "C" int linompss_dltrsm(char uplo, char trans, char diag, int bm, int bn, int m, int n, double alpha,
                               const double *A, int lda, double *B, int ldb)
{
    return host_ltrsm<double>(uplo, trans, diag, bm, bn, m, n, alpha, A, lda, B, ldb);
}
extern "C" int linompss_sltrsm(char uplo, char trans, char diag, int bm, int bn, int m, int n, float alpha,
                               const float *A, int lda, float *B, int ldb)
{
    return host_ltrsm<float>(uplo, trans, diag, bm, bn, m, n, alpha, A, lda, B, ldb);
}

// ---------------------------------------------------------------------------
// Plan introspection (host only; no device needed).

static int plan_host(int what, int n, int b, int nrhs, const linompss_exec_opts *o, lx_planner &pl)
{
    if (n <= 0 || b <= 0 || nrhs < 0) return -1;
    lx_shape sh = {what, 0, n, b, nrhs};
    try {
        if (!build_plan(pl, sh)) return -1;
    } catch (...) {
        return -1;
    }
    (void)o;
    return 0;
}

extern "C" int linompss_plan_tdg(int what, int n, int b, int nrhs, int *types, int *coords, int max_tasks,
                                 int *edges, int max_edges, int *ntasks, int *nedges)
{
    lx_plan_params P;
    P.unit_m = LX_UNIT_M;
    P.unit_n = LX_UNIT_N;
    lx_planner pl(P);
    int w = what == 2 ? LX_W_POSV_RL : what;
    if (plan_host(w, n, b, std::max(nrhs, what == 2 ? 1 : 0), nullptr, pl)) return -2;
    const auto &T = pl.tasks();
    int nt = (int)T.size(), ne = 0;
    for (const auto &t : T) ne += (int)t.preds.size();
    if (ntasks) *ntasks = nt;
    if (nedges) *nedges = ne;
    if (nt > max_tasks || ne > max_edges) return -1;
    int e = 0;
    for (int q = 0; q < nt; ++q) {
        if (types) types[q] = T[q].type;
        if (coords) {
            coords[3 * q] = T[q].ti;
            coords[3 * q + 1] = T[q].tj;
            coords[3 * q + 2] = T[q].tk;
        }
        for (int p : T[q].preds) {
            if (edges) {
                edges[2 * e] = p;
                edges[2 * e + 1] = q;
            }
            ++e;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Recorded factor configurations (the search itself is in lx_tune.hip).

static bool cfg_valid(const linompss_factor_config *c)
{
    return c->b > 0 && c->potrf_sub >= 64 && c->potrf_sub % 64 == 0 && c->gemm_wide_m >= 128 &&
           c->gemm_wide_m % 128 == 0 && c->gemm_wide_n >= 128 && c->gemm_wide_n % 128 == 0 &&
           c->trsm_rows_wide >= 64 && c->trsm_rows_wide % 64 == 0 && c->wide_tail >= -1 &&
           c->update_group >= 1 && c->update_group <= 8;
}

extern "C" int linompss_set_factor_config(int dtype, int variant, int n, int update_prec,
                                          const linompss_factor_config *cfg)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (variant != 0 && variant != 1) return -2;
    if (n <= 0) return -3;
    if (update_prec < 0 || update_prec > 2 || (dtype == 0 && update_prec != 0)) return -4;
    if (!cfg || !cfg_valid(cfg)) return -5;
    std::lock_guard<std::mutex> g(g_mu);
    for (auto &e : g_cfgs)
        if (e.dtype == dtype && e.variant == variant && e.n == n && e.b == cfg->b && e.update_prec == update_prec) {
            e.cfg = *cfg;
            return 0;
        }
    g_cfgs.push_back({dtype, variant, n, cfg->b, update_prec, *cfg});
    return 0;
}

extern "C" int linompss_get_factor_config(int dtype, int variant, int n, int b, int update_prec,
                                          linompss_factor_config *cfg)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (variant != 0 && variant != 1) return -2;
    if (n <= 0) return -3;
    if (b <= 0) return -4;
    if (update_prec < 0 || update_prec > 2 || (dtype == 0 && update_prec != 0)) return -5;
    if (!cfg) return -6;
    std::lock_guard<std::mutex> g(g_mu);
    if (const lx_cfg_entry *e = cfg_find(dtype, variant, n, b, update_prec)) {
        *cfg = e->cfg;
        return 1;
    }
    // the defaults (environment knobs included)
    lx_cfg_defaults(dtype, update_prec, b, cfg);
    return 0;
}

bool lx_cfg_lookup(int dtype, int variant, int n, int b, int update_prec, linompss_factor_config *cfg)
{
    std::lock_guard<std::mutex> g(g_mu);
    const lx_cfg_entry *e = cfg_find(dtype, variant, n, b, update_prec);
    if (e) *cfg = e->cfg;
    return e != nullptr;
}

void lx_cfg_forget(int dtype, int variant, int n, int b, int update_prec)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_cfgs.erase(std::remove_if(g_cfgs.begin(), g_cfgs.end(),
                                [&](const lx_cfg_entry &e) {
                                    return e.dtype == dtype && e.variant == variant && e.n == n && e.b == b &&
                                           e.update_prec == update_prec;
                                }),
                 g_cfgs.end());
}

void lx_cfg_defaults(int dtype, int update_prec, int b, linompss_factor_config *cfg)
{
    linompss_exec_opts o = {};
    o.policy = 1;
    o.update_prec = update_prec;
    lx_plan_params P;
    fill_params(P, dtype, 512, &o, nullptr, b);
    P.update_group = default_update_group(b);
    if (const char *ev = getenv("LINOMPSS_UPDATE_GROUP")) {
        const int v = atoi(ev);
        if (v >= 1 && v <= 8) P.update_group = v;
    }
    *cfg = {b, P.potrf_sub, P.gemm_wide_m, P.gemm_wide_n, P.trsm_rows_wide, P.wide_tail, P.update_group, 0.0};
}

extern "C" int linompss_clear_factor_configs(void)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_cfgs.clear();
    return 0;
}

extern "C" int linompss_plan_info(int what, int n, int b, int nrhs, const linompss_exec_opts *opts, int *units,
                                  int *tasks, double *sim_ms)
{
    lx_plan_params P;
    {
        std::lock_guard<std::mutex> g(g_mu);  // the recorded configurations
        const lx_shape sh = {what, 0, n, b, nrhs};
        fill_params(P, 0, 512, opts, &sh);
    }
    lx_planner pl(P);
    if (plan_host(what, n, b, nrhs, opts, pl)) return -2;
    if (units) *units = (int)pl.units().size();
    if (tasks) *tasks = (int)pl.tasks().size();
    if (sim_ms) *sim_ms = pl.simulated_time() * 1e3;
    return 0;
}

extern "C" const char *linompss_version(void) { return "linompss-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Device-memory plumbing.

extern "C" int linompss_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" void *linompss_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

extern "C" int linompss_dev_free(void *p) { return hipFree(p) == hipSuccess ? 0 : LINOMPSS_ERR_HIP; }

// Device-to-device copies run as this kernel (16-byte loads and stores when
// both pointers and the size allow it): a plain copy of known size whose HBM
// counters calibrate the traffic figures of the other kernels.
__global__ void __launch_bounds__(256) lx_copy_kernel(char *__restrict__ dst, const char *__restrict__ src, size_t bytes)
{
    const size_t stride = (size_t)gridDim.x * 256;
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0) {
        const size_t nv = bytes / 16;
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nv; i += stride)
            ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
    } else {
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < bytes; i += stride) dst[i] = src[i];
    }
}

extern "C" int linompss_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
    if (kind < 1 || kind > 3) return -4;
    if (bytes == 0) return 0;
    if (kind == 3) {
        hipLaunchKernelGGL(lx_copy_kernel, dim3(4096), dim3(256), 0, nullptr, (char *)dst, (const char *)src, bytes);
        if (hipGetLastError() != hipSuccess) return LINOMPSS_ERR_HIP;
        return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
    }
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
}

extern "C" int linompss_dev_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP; }
