// lx_planner.hpp -- host-side construction of execution plans (C++).
//
// A plan is the reference's task graph made explicit: the same loop nests the
// reference's master thread walks to spawn OmpSs tasks (CHOL_RL
// chol_rlmain.c:25-50, CHOL_LL chol_llmain.c:26-53, the lower branches of
// bblas_dltrsm bblas_trsm.h:119-160), the same tile operations, and the same
// RAW / WAW dependency semantics that Nanos++ derives from the tiles' first
// elements.  Each task is cut into work units for the device; the units are
// then ordered by a list-scheduling simulation (the "priority queue" of the
// reference's `priority(p)` clauses, evaluated once on the host instead of by
// a runtime on every spawn).
#pragma once
#include <stdint.h>
#include <vector>

#include "lx_plan.hpp"

// True when each unit's operands lie inside single b x b tiles (the
// condition for running a plan on a hypermatrix).
bool lx_units_in_tiles(const std::vector<lx_unit> &U, int b);

struct lx_plan_params {
    int unit_m = 128;       // GEMM / SYRK unit rows (one workgroup tile)
    int unit_n = 128;       // GEMM / SYRK unit cols
    int gemm_wide_n = 512;  // GEMM unit cols off the lookahead column (one fence per 128 x 512)
    int gemm_wide_m = 128;  // max GEMM unit rows off the lookahead column (128-row blocks, one fence per unit)
    double wide_units = 2.0;  // taller units only while a step still has wide_units units per workgroup
    int wide_tail = 0;      // the last wide_tail panels (-1: half of them) use unit_n-column GEMM units throughout
    int potrf_sub = 128;    // diagonal tiles wider than this factor as a sub-DAG of sub-tiles of this size
    int update_group = 1;   // RL: panels per trailing-update pass for tiles right of a group of that
                            // many steps (1: one pass per panel; g: K = g b for most of the matrix)
    int update_prec = 0;    // trailing GEMM/SYRK units of the tile loop: 1 split-BF16, 2 split-FP16 matrix cores
    int bx_wide = 1;        // split-BF16 units: 128 x 256 blocks (0: 128 x 128 blocks only)
    int trsm_rows = 64;     // panel TRSM rows per unit
    int trsm_rows_wide = 64;  // panel TRSM rows per unit off the lookahead tile (64-row passes, one fence)
    int gemv_rows = 128;    // RHS GEMV output rows per unit
    int policy = 1;         // 0: reference priorities + FIFO, 1: critical-path (bottom level)
    int workers = 512;      // workgroups assumed resident in the simulation
    double flop_rate = 1.5e11;   // per-workgroup MFMA rate (flop/s) for GEMM/SYRK
    double potrf_rate = 3.0e10;  // per-workgroup rate of the diagonal factorisation
    double trsm_rate = 6.0e10;   // per-workgroup rate of the panel solve
    double byte_rate = 4.0e10;   // per-workgroup streaming rate of the RHS sweeps (B/s)
    double unit_overhead = 3e-6; // grab + wait + publish (s)
    int elem_bytes = 8;
};

// Task types, as in the reference's tile tasks.
enum lx_task_type : int32_t {
    LX_T_POTRF = 0,
    LX_T_TRSM = 1,
    LX_T_GEMM = 2,
    LX_T_SYRK = 3,
    LX_T_SV_TRSM = 4,   // sweep TRSM on a right-hand-side block
    LX_T_SV_GEMM = 5    // sweep GEMM on a right-hand-side block
};

struct lx_task {
    int32_t type;
    int32_t sys;
    int32_t ti, tj, tk;     // tile coordinates (meaning per type)
    int32_t prio;           // reference priority (9999 / 9 / 1)
    int32_t first_unit, nunits;
    double unit_cost;       // simulated seconds per unit
    std::vector<int32_t> preds;  // task-level predecessors (last writers)
};

class lx_planner {
  public:
    explicit lx_planner(const lx_plan_params &p) : P(p) {}

    // Registers a system; returns its index.  Counters for its tiles and
    // right-hand-side blocks are allocated here.
    int add_system(int n, int b, int nrhs);

    // Reference loop nests (see file header).
    void factor_rl(int sys);
    void factor_ll(int sys);
    void sweep_forward(int sys);   // L Y = X
    void sweep_backward(int sys);  // L^T Z = Y

    // Orders units by simulation and returns the device unit array.
    void finalize();

    const std::vector<lx_unit> &units() const { return U; }
    const std::vector<lx_task> &tasks() const { return Tk; }
    int num_counters() const { return nctr; }
    double simulated_time() const { return sim_time; }

  private:
    struct sysinfo {
        int n, b, mt, nrhs;
        int ctr_base;   // tiles: ctr_base + i*(i+1)/2 + j ; rhs blocks: ctr_base + mt*(mt+1)/2 + i
    };
    int tile_ctr(int s, int i, int j) const { return S[s].ctr_base + i * (i + 1) / 2 + j; }
    int rhs_ctr(int s, int i) const { return S[s].ctr_base + S[s].mt * (S[s].mt + 1) / 2 + i; }
    int new_task(int type, int sys, int ti, int tj, int tk, int prio, int out_ctr, const int *in_ctrs, int nin);
    void add_unit(int task, const lx_unit &proto);
    // A unit with its own dependency list (units inside one task that wait on
    // each other, created in a topological order).
    void add_unit_deps(int task, const lx_unit &proto, const std::vector<std::pair<int32_t, int32_t>> &deps);
    int alloc_counters(int count);
    // aux0 flag of the tile loop's GEMM/SYRK units: 2 = split-BF16 matrix cores
    int upd_flag() const { return P.update_prec >= 1 ? (P.bx_wide ? 6 : 2) | (P.update_prec == 2 ? 8 : 0) : 0; }
    // Units of the diagonal-tile factorisation POTRF(k) of task t (its output
    // counter outc, global column offset k*b, width bs).
    void potrf_units(int t, int s, int kb, int bs, int outc);

    lx_plan_params P;
    std::vector<sysinfo> S;
    std::vector<lx_task> Tk;
    std::vector<lx_unit> U;           // creation order until finalize(), then execution order
    std::vector<int32_t> issued;      // units issued per counter so far
    std::vector<int32_t> last_writer; // task that last wrote each counter's tile
    std::vector<int32_t> unit_task;   // creation-order unit -> task
    // per task: dependency list applied to all of its units
    std::vector<std::vector<std::pair<int32_t, int32_t>>> task_deps;
    int nctr = 0;
    double sim_time = 0.0;
    bool finalized = false;
};
