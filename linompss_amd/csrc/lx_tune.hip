// lx_tune.hip -- the factor autotuner: chooses the tile size and the work-unit
// shapes of the persistent executor's plans by timing them on the device.
//
// The reference fixes its task shapes in code (one task per b x b tile,
// chol_rlmain.c:25-50) and takes b from its caller.  Here a tile task is cut
// into work units whose shapes (diagonal sub-tiles, trailing-update unit
// heights and widths, panel-solve heights, narrow units in the last panels)
// decide how well the persistent launch fills the chip (and how many panels'
// updates a tile right of the diagonal chain takes per pass); the best shapes
// depend on n, b, the precision and the update arithmetic.  The search is a
// coordinate descent from the library's defaults over a small grid per knob:
// each knob in turn takes each of its values with the others held at the
// best so far, and keeps the fastest; when the tile size is searched too,
// every candidate tile is timed with its defaults and the descent runs from
// the two fastest.  A candidate's time is the fastest of
// `reps` factor launches (HIP events on the launch stream) of one generated
// matrix, restored from a device copy before every launch.  The winner is
// recorded for its shape, so every later factor of that shape runs with it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "linompss.h"
#include "lx_internal.hpp"

namespace {

// a_ii = 1 and a_ij = a_ji in [0, 1/n) (the generator filled U[0, 1/n)):
// strictly diagonally dominant, so SPD, and every entry of its factor lies in
// [0, 1], which every update precision accepts.
template <typename T>
__global__ void lx_tune_diag_kernel(T *A, int64_t lda, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(int64_t)i * lda + i] = T(1);
}

struct tune_ctx {
    int dtype, variant, n, update_prec, reps;
    void *dA = nullptr, *dW = nullptr;
    size_t bytes = 0;
    std::vector<linompss_factor_config> tried;
};

// The fastest of t.reps factor launches with cfg recorded for its shape;
// 0 or an error code in *err.
double time_cfg(tune_ctx &t, linompss_factor_config cfg, int *err)
{
    *err = linompss_set_factor_config(t.dtype, t.variant, t.n, t.update_prec, &cfg);
    if (*err) return 0.0;
    linompss_exec_opts o = {};
    o.policy = 1;
    o.update_prec = t.update_prec;
    double best = 1e30;
    for (int r = 0; r < t.reps; ++r) {
        if (hipMemcpy(t.dW, t.dA, t.bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
            *err = LINOMPSS_ERR_HIP;
            return 0.0;
        }
        const int rc = linompss_chol_dev(t.dtype, t.variant, t.n, cfg.b, t.dW, t.n, nullptr, &o);
        if (rc != 0) {  // the matrix is SPD: an info > 0 is a failure of the launch
            *err = rc > 0 ? LINOMPSS_ERR_SCHED : rc;
            return 0.0;
        }
        linompss_exec_stats st = {};
        linompss_last_exec_stats(&st);
        best = std::min(best, st.exec_ms);
    }
    cfg.ms = best;
    t.tried.push_back(cfg);
    return best;
}

}  // namespace

extern "C" int linompss_tune_factor(int dtype, int variant, int n, int b, int update_prec, int reps,
                                    linompss_factor_config *best, linompss_factor_config *tried, int max_tried,
                                    int *ntried)
{
    if (dtype != 0 && dtype != 1) return -1;
    if (variant != 0 && variant != 1) return -2;
    if (n <= 0) return -3;
    if (b < 0) return -4;
    if (update_prec < 0 || update_prec > 2 || (dtype == 0 && update_prec != 0)) return -5;
    if (!best) return -7;
    if (max_tried < 0 || (max_tried > 0 && !tried)) return -9;
    tune_ctx t;
    t.dtype = dtype;
    t.variant = variant;
    t.n = n;
    t.update_prec = update_prec;
    t.reps = reps > 0 ? reps : 3;

    // tile sizes searched, and what was recorded for each before (restored
    // for every tile size that does not win, and for all on failure)
    std::vector<int> tiles;
    if (b > 0) {
        tiles.push_back(b);
    } else {
        // 128 and 256 only up to n = 16384: below that order the diagonal
        // chain, not the trailing updates, sets the factor's time (n = 8192,
        // FP64: b = 128 11.7 ms, 256 12.1 ms, 384 12.6 ms with their defaults)
        for (int c : {128, 256, 384, 512, 768, 1024})
            if (c <= n && (c >= 384 || n <= 16384)) tiles.push_back(c);
        if (tiles.empty()) tiles.push_back(n);
    }
    std::vector<std::pair<bool, linompss_factor_config>> prior(tiles.size());
    for (size_t q = 0; q < tiles.size(); ++q)
        prior[q].first = lx_cfg_lookup(dtype, variant, n, tiles[q], update_prec, &prior[q].second);

    const size_t es = dtype == 0 ? sizeof(double) : sizeof(float);
    t.bytes = (size_t)n * n * es;
    if (hipMalloc(&t.dA, t.bytes) != hipSuccess) return LINOMPSS_ERR_NOMEM;
    if (hipMalloc(&t.dW, t.bytes) != hipSuccess) {
        (void)hipFree(t.dA);
        return LINOMPSS_ERR_NOMEM;
    }
    int rc = linompss_gen_uniform(dtype, n, n, 20260u, 0.0, 1.0 / n, t.dA, n);
    if (rc == 0) {
        if (dtype == 0)
            hipLaunchKernelGGL(lx_tune_diag_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, 0, (double *)t.dA,
                               (int64_t)n, n);
        else
            hipLaunchKernelGGL(lx_tune_diag_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, 0, (float *)t.dA,
                               (int64_t)n, n);
        rc = hipDeviceSynchronize() == hipSuccess ? 0 : LINOMPSS_ERR_HIP;
    }

    // one pass of coordinate descent over the unit shapes for a fixed tile,
    // from that tile's defaults (timed as t0)
    struct knob {
        int linompss_factor_config::*f;
        std::vector<int> v;
    };
    const std::vector<knob> knobs = {
        {&linompss_factor_config::update_group, {1, 2, 3, 4, 6, 8}},
        {&linompss_factor_config::potrf_sub, {128, 256, 512}},
        {&linompss_factor_config::gemm_wide_m, {128, 256, 512}},
        {&linompss_factor_config::gemm_wide_n, {256, 512, 1024}},
        {&linompss_factor_config::trsm_rows_wide, {64, 128, 256}},
        {&linompss_factor_config::wide_tail, {0, 8, 16, -1}},
    };
    auto descend = [&](linompss_factor_config &cur, double &tbest) {
        for (size_t q = 0; q < knobs.size() && rc == 0; ++q) {
            const linompss_factor_config base = cur;
            for (int v : knobs[q].v) {
                linompss_factor_config c = base;
                if (c.*knobs[q].f == v) continue;
                c.*knobs[q].f = v;
                if (knobs[q].f == &linompss_factor_config::potrf_sub && v >= c.b && base.potrf_sub >= c.b)
                    continue;  // no sub-DAG either way
                const double ms = time_cfg(t, c, &rc);
                if (rc) break;
                if (ms < tbest) {
                    tbest = ms;
                    cur = c;
                }
            }
        }
    };
    // every candidate tile with its defaults, then the descent from the two
    // fastest (a tile's best unit shapes differ from another's, so the tile is
    // not fixed before the shapes are searched)
    std::vector<std::pair<double, linompss_factor_config>> start;
    for (size_t q = 0; q < tiles.size() && rc == 0; ++q) {
        linompss_factor_config c = {};
        lx_cfg_defaults(dtype, update_prec, tiles[q], &c);
        const double ms = time_cfg(t, c, &rc);
        if (rc == 0) start.push_back({ms, c});
    }
    std::sort(start.begin(), start.end(), [](const auto &a, const auto &b2) { return a.first < b2.first; });
    linompss_factor_config cur = {};
    double tbest = 1e30;
    for (size_t q = 0; q < start.size() && q < 2 && rc == 0; ++q) {
        linompss_factor_config c = start[q].second;
        double tb = start[q].first;
        descend(c, tb);
        if (rc == 0 && tb < tbest) {
            tbest = tb;
            cur = c;
        }
    }

    // put back what other tile sizes had; record the winner
    for (size_t q = 0; q < tiles.size(); ++q) {
        if (rc == 0 && tiles[q] == cur.b) continue;
        if (prior[q].first)
            (void)linompss_set_factor_config(dtype, variant, n, update_prec, &prior[q].second);
        else
            lx_cfg_forget(dtype, variant, n, tiles[q], update_prec);
    }
    if (rc == 0) {
        cur.ms = tbest;
        rc = linompss_set_factor_config(dtype, variant, n, update_prec, &cur);
        *best = cur;
    }
    if (ntried) *ntried = (int)t.tried.size();
    for (int q = 0; q < max_tried && q < (int)t.tried.size(); ++q) tried[q] = t.tried[q];
    (void)hipFree(t.dA);
    (void)hipFree(t.dW);
    return rc;
}
