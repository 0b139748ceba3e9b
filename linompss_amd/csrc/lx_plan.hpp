// lx_plan.hpp -- task-graph plan shared by the host planner and the
// persistent device executor.
//
// The reference builds its task graph at run time: the master thread walks
// the RL / LL loop nests (src/cholesky/dense/chol_rlmain.c:25-50,
// chol_llmain.c:26-53, src/blocked_blas/bblas_trsm.h:119-160) and Nanos++
// derives RAW / WAW edges from the first element of each tile (in/inout
// clauses, e.g. tasks_potrf.h:19).  Here the host planner walks the same loop
// nests once per shape, splits each tile task into work units, and records for
// every unit the (counter, threshold) pairs it must see before it may run.
// Counters live in device memory, one per tile (and one per right-hand-side
// block); a unit bumps its tile's counter by one when it completes.  The
// device grabs units in the planner's priority order, so no ready queue is
// needed on the device and every wait is on an earlier unit (deadlock free
// whatever the number of resident workgroups).
#pragma once
#include <stdint.h>

enum lx_op : int32_t {
    LX_OP_POTRF = 0,      // C tile (M x M) := chol(C), lower
    LX_OP_TRSM_RLTN = 1,  // C rows (M x N) := C * L^{-T}, L = A (N x N lower)
    LX_OP_SYRK = 2,       // C (M x N sub-block of a diagonal tile) -= A B^T, lower mask
    LX_OP_GEMM_NT = 3,    // C (M x N) -= A (M x K) * B (N x K)^T
    LX_OP_TRSV = 4,       // X block (M x nrhs) := L^{-1} X or L^{-T} X (aux0 = trans)
    LX_OP_GEMV_RHS = 5,   // X_c (M x nrhs) -= op(L) X_b (aux0 = trans; K inner)
    LX_OP_NOP = 6
};

// Operand base selectors (which system buffer an offset refers to).
enum lx_base : int32_t { LX_BASE_A = 0, LX_BASE_X = 1 };

#define LX_MAX_DEPS 3

struct lx_unit {
    int32_t op;
    int32_t sys;           // system index (batched plans hold many systems)
    int32_t M, N, K;       // unit extents
    int32_t ctr;           // counter bumped on completion
    int32_t nd;            // number of dependencies
    int32_t dc[LX_MAX_DEPS];
    int32_t dt[LX_MAX_DEPS];
    int32_t aux0;          // op specific: trans flag / lower flag
    int32_t aux1;          // op specific: gdiag (SYRK) / global column offset (POTRF)
    int32_t cbase, abase, bbase;
    int64_t c, a, b;       // element offsets of C, A, B in their base buffers
};
static_assert(sizeof(lx_unit) == 96, "lx_unit layout");

struct lx_sys {
    void *A;               // matrix (T*), column-major
    void *X;               // right-hand sides (T*), column-major, may be null
    int64_t lda, ldx;
    int32_t n, nrhs;
    int32_t info_slot;     // index in the info array
    int32_t pad_;
    void *W;               // factor workspace (T*), column-major (n + 64) x 64:
                           // inverses of the 64 x 64 diagonal blocks, rows as in A
    int64_t ldw;
    void *const *tiles;    // hypermatrix A (factor plans only): element (r, c) in
                           // tile tiles[(c / tb) * tmt + r / tb] at (c % tb) * tb
                           // + r % tb; null for a column-major A (lda)
    int32_t tmt, tb;
    void *planes;          // FP32 factor with split-FP16 updates (else null): the
                           // FP16 planes hi, lo of 2^12 L, written once by the
                           // unit that produces each element of L (TRSM units) and
                           // read by the trailing GEMM / SYRK units.  Layout
                           // [plane][slab c/16][row][16]: the 16 columns of a slab
                           // of one row are 32 contiguous bytes, so 32 rows of a
                           // slab (one MFMA operand tile) are 1 KB contiguous
    int64_t ldp;           // rows per slab (>= n + 256, a multiple of 32)
    int32_t nsig, pad2_;   // slabs per plane (ceil(n / 16))
};

// Device-side status words (one per launch).
struct lx_status {
    int32_t head;          // next unit index to grab
    int32_t abort;         // nonzero once a wait timed out
    int32_t timeouts;      // units that gave up waiting
    int32_t done;          // units completed
};
