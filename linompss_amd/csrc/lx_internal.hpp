// lx_internal.hpp -- host-side helpers shared by the library's translation
// units (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "lx_plan.hpp"
#include "linompss.h"

// Plan kinds.
enum lx_what : int {
    LX_W_FACTOR_RL = 0,
    LX_W_FACTOR_LL = 1,
    LX_W_POSV_RL = 2,    // RL factor + forward + backward sweep
    LX_W_SWEEPS = 3,     // forward + backward sweep on a factored matrix
    LX_W_FWD = 4,
    LX_W_BWD = 5
};

struct lx_shape {
    int what, dtype, n, b, nrhs;
    int tiled = 0;  // a hypermatrix factor: every unit's operands inside one b x b tile
};

// Runs one plan over nsys systems sharing the same shape family (the plan was
// built for exactly these systems).  info_out[s] receives each system's info.
// Returns 0 or a LINOMPSS_ERR_* code.
// dtiles (factor plans only): a device table of tile pointers for a
// hypermatrix A (tile (i, j) at dtiles[j * (n / b) + i], b x b, leading
// dimension b); dA is then unused.
int lx_run_single(const lx_shape &sh, void *dA, int64_t lda, void *dX, int64_t ldx, hipStream_t stream,
                  const linompss_exec_opts *opts, int *info_out, void *const *dtiles = nullptr);

// Triangular sweeps with dedicated kernels (lx_sweep.hip): which = 1 forward
// (L Y = X), 2 backward (L^T Z = Y), 3 both.  Synchronises the stream; returns
// LINOMPSS_ERR_SCHED if a wait inside a sweep timed out.
int lx_sweeps_dev(int dtype, int n, int nrhs, const void *L, int64_t ldl, void *X, int64_t ldx, int which,
                  hipStream_t s);

// 128-row FP32 sweeps (lx_sweep.hip): whether a solve of this shape takes
// them (LINOMPSS_SWEEP2_MIN_N: the smallest order, 0 = never; default 2048),
// the bytes of their step matrices, the pre-pass that writes them from L
// (enqueued only), and the sweeps over them (which as above; synchronises the
// stream).  The caller holds lx_dev_mutex().
int lx_sw2_enabled(int dtype, int n, int nrhs);
size_t lx_sw2_prep_bytes(int n);
int lx_sw2_prep(int n, const float *L, int64_t ldl, float *prep, hipStream_t s);
int lx_sw2_sweeps(int n, int nrhs, const float *L, int64_t ldl, float *X, int64_t ldx, int which, const float *prep,
                  hipStream_t s);

// The lock that serialises the executor's plans, its scratch slots and the
// sweep kernels on a device.
std::mutex &lx_dev_mutex();

// Grow-only device scratch owned by the library (one per slot and device).
void *lx_scratch(int slot, size_t bytes);

// Recorded factor configurations (lx_exec.hip): lookup (false when the shape
// has none), removal, and the library's defaults for a tile size.
bool lx_cfg_lookup(int dtype, int variant, int n, int b, int update_prec, linompss_factor_config *cfg);
void lx_cfg_forget(int dtype, int variant, int n, int b, int update_prec);
void lx_cfg_defaults(int dtype, int update_prec, int b, linompss_factor_config *cfg);

// Elementwise helpers enqueued on a stream.
void lx_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s);

#define LX_HIP_TRY(expr)                                                                                    \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return LINOMPSS_ERR_HIP;                                                      \
    } while (0)
