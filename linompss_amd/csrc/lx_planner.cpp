// lx_planner.cpp -- see lx_planner.hpp.
#include "lx_planner.hpp"

#include <algorithm>
#include <functional>
#include <queue>
#include <stdexcept>

// Reference task priorities: chol_rlmain.c:16-19 (RL), chol_llmain.c:17-20
// (LL); the sweep tasks of bblas_trsm.h:119-160 pass priority 1.
static const int PRIO_HI = 9999, PRIO_LO = 9, PRIO_SWEEP = 1;

int lx_planner::add_system(int n, int b, int nrhs)
{
    if (finalized) throw std::logic_error("planner finalized");
    sysinfo s;
    s.n = n;
    s.b = b;
    s.mt = (n + b - 1) / b;
    s.nrhs = nrhs;
    s.ctr_base = nctr;
    nctr += s.mt * (s.mt + 1) / 2 + s.mt;
    issued.resize(nctr, 0);
    last_writer.resize(nctr, -1);
    S.push_back(s);
    return (int)S.size() - 1;
}

int lx_planner::new_task(int type, int sys, int ti, int tj, int tk, int prio, int out_ctr, const int *in_ctrs,
                         int nin)
{
    lx_task t;
    t.type = type;
    t.sys = sys;
    t.ti = ti;
    t.tj = tj;
    t.tk = tk;
    t.prio = prio;
    t.first_unit = (int)U.size();
    t.nunits = 0;
    t.unit_cost = 0.0;
    std::vector<std::pair<int32_t, int32_t>> deps;
    // WAW on the output tile: every earlier writer must be complete.
    if (issued[out_ctr] > 0) {
        deps.push_back({out_ctr, issued[out_ctr]});
        t.preds.push_back(last_writer[out_ctr]);
    }
    // RAW on each input tile.
    for (int q = 0; q < nin; ++q) {
        int c = in_ctrs[q];
        if (issued[c] > 0) {
            deps.push_back({c, issued[c]});
            if (std::find(t.preds.begin(), t.preds.end(), last_writer[c]) == t.preds.end())
                t.preds.push_back(last_writer[c]);
        }
    }
    if ((int)deps.size() > LX_MAX_DEPS) throw std::logic_error("too many dependencies");
    Tk.push_back(t);
    task_deps.push_back(deps);
    return (int)Tk.size() - 1;
}

void lx_planner::add_unit(int task, const lx_unit &proto)
{
    lx_unit u = proto;
    const auto &deps = task_deps[task];
    u.nd = (int)deps.size();
    for (int d = 0; d < LX_MAX_DEPS; ++d) {
        u.dc[d] = d < u.nd ? deps[d].first : 0;
        u.dt[d] = d < u.nd ? deps[d].second : 0;
    }
    U.push_back(u);
    unit_task.push_back(task);
    Tk[task].nunits++;
}

void lx_planner::add_unit_deps(int task, const lx_unit &proto, const std::vector<std::pair<int32_t, int32_t>> &deps)
{
    if ((int)deps.size() > LX_MAX_DEPS) throw std::logic_error("too many dependencies");
    lx_unit u = proto;
    u.nd = (int)deps.size();
    for (int d = 0; d < LX_MAX_DEPS; ++d) {
        u.dc[d] = d < u.nd ? deps[d].first : 0;
        u.dt[d] = d < u.nd ? deps[d].second : 0;
    }
    U.push_back(u);
    unit_task.push_back(task);
    Tk[task].nunits++;
}

int lx_planner::alloc_counters(int count)
{
    const int base = nctr;
    nctr += count;
    issued.resize(nctr, 0);
    last_writer.resize(nctr, -1);
    return base;
}

static lx_unit proto_unit(int op, int sys)
{
    lx_unit u = {};
    u.op = op;
    u.sys = sys;
    u.cbase = u.abase = u.bbase = LX_BASE_A;
    return u;
}

// Element offset of (row, col) in a column-major matrix with leading dim ld
// is computed on the device (ld is per system); the plan stores (row, col)
// packed as row + col * 2^32 so that one plan serves any lda.
static inline int64_t rc(int64_t r, int64_t c) { return (c << 32) | r; }

// POTRF(k): one unit when the tile is at most potrf_sub wide; otherwise the
// tile is factored as a right-looking sub-DAG of potrf_sub-wide sub-tiles --
// POTRF, TRSM and SYRK/GEMM units with counters of their own -- so that the
// diagonal factorisation, the step every other task of the next panel waits
// for, is spread over several workgroups.  Every sub-unit waits (directly or
// through the first sub-POTRF) for the tile's earlier updates; only the last
// sub-POTRF bumps the tile's counter, and it runs after every other sub-unit
// (each one precedes it in the sub-DAG), so that bump publishes the whole tile.
void lx_planner::potrf_units(int t, int s, int kb, int bs, int outc)
{
    const std::vector<std::pair<int32_t, int32_t>> tile_deps = task_deps[t];
    const int sb = P.potrf_sub;
    if (bs <= sb || sb <= 0) {
        lx_unit u = proto_unit(LX_OP_POTRF, s);
        u.M = u.N = u.K = bs;
        u.ctr = outc;
        u.c = rc(kb, kb);
        u.aux1 = kb;
        Tk[t].unit_cost = (double)bs * bs * bs / 3.0 / P.potrf_rate + P.unit_overhead;
        add_unit_deps(t, u, tile_deps);
        return;
    }
    const int ns = (bs + sb - 1) / sb;
    const int base = alloc_counters(ns * (ns + 1) / 2);
    auto sub = [&](int i, int j) { return base + i * (i + 1) / 2 + j; };
    std::vector<int32_t> cnt(ns * (ns + 1) / 2, 0);  // units issued per sub-counter
    auto cnt_of = [&](int i, int j) -> int32_t & { return cnt[i * (i + 1) / 2 + j]; };
    for (int kk = 0; kk < ns; ++kk) {
        const int k0 = kb + kk * sb, kbs = std::min(sb, bs - kk * sb);
        {   // sub-POTRF(kk)
            std::vector<std::pair<int32_t, int32_t>> d = kk == 0 ? tile_deps : std::vector<std::pair<int32_t, int32_t>>();
            if (cnt_of(kk, kk) > 0) d.push_back({sub(kk, kk), cnt_of(kk, kk)});
            lx_unit u = proto_unit(LX_OP_POTRF, s);
            u.M = u.N = u.K = kbs;
            const bool last = kk == ns - 1;
            u.ctr = last ? outc : sub(kk, kk);
            u.c = rc(k0, k0);
            u.aux1 = k0;
            add_unit_deps(t, u, d);
            if (!last) cnt_of(kk, kk) += 1;
        }
        for (int ii = kk + 1; ii < ns; ++ii) {  // sub-TRSM(ii, kk)
            const int i0 = kb + ii * sb, ibs = std::min(sb, bs - ii * sb);
            std::vector<std::pair<int32_t, int32_t>> d = {{sub(kk, kk), cnt_of(kk, kk)}};
            if (cnt_of(ii, kk) > 0) d.push_back({sub(ii, kk), cnt_of(ii, kk)});
            int nu = 0;
            for (int r0 = 0; r0 < ibs; r0 += P.trsm_rows, ++nu) {
                lx_unit u = proto_unit(LX_OP_TRSM_RLTN, s);
                u.M = std::min(P.trsm_rows, ibs - r0);
                u.N = kbs;
                u.ctr = sub(ii, kk);
                u.c = rc(i0 + r0, k0);
                u.a = rc(k0, k0);
                add_unit_deps(t, u, d);
            }
            cnt_of(ii, kk) += nu;
        }
        for (int ii = kk + 1; ii < ns; ++ii) {  // sub-SYRK(ii, kk) and sub-GEMM(ii, jj, kk)
            const int i0 = kb + ii * sb, ibs = std::min(sb, bs - ii * sb);
            for (int jj = kk + 1; jj <= ii; ++jj) {
                const int j0 = kb + jj * sb, jbs = std::min(sb, bs - jj * sb);
                std::vector<std::pair<int32_t, int32_t>> d = {{sub(ii, kk), cnt_of(ii, kk)}};
                if (jj != ii) d.push_back({sub(jj, kk), cnt_of(jj, kk)});
                if (cnt_of(ii, jj) > 0) d.push_back({sub(ii, jj), cnt_of(ii, jj)});
                int nu = 0;
                for (int c0 = 0; c0 < jbs; c0 += P.unit_n)
                    for (int r0 = 0; r0 < ibs; r0 += P.unit_m) {
                        if (ii == jj && r0 + P.unit_m - 1 < c0) continue;  // strictly upper
                        lx_unit u = proto_unit(ii == jj ? LX_OP_SYRK : LX_OP_GEMM_NT, s);
                        u.M = std::min(P.unit_m, ibs - r0);
                        u.N = std::min(P.unit_n, jbs - c0);
                        u.K = kbs;
                        u.ctr = sub(ii, jj);
                        u.c = rc(i0 + r0, j0 + c0);
                        u.a = rc(i0 + r0, k0);
                        u.b = rc(j0 + c0, k0);
                        if (ii == jj) {
                            u.aux0 = 1;
                            u.aux1 = r0 - c0;
                        }
                        add_unit_deps(t, u, d);
                        ++nu;
                    }
                cnt_of(ii, jj) += nu;
            }
        }
    }
    // simulated cost per unit: the sub-DAG's chain spread over its units
    const double chain = ns * ((double)sb * sb * sb / 3.0 / P.potrf_rate + (double)sb * sb * sb / P.trsm_rate +
                               2.0 * sb * sb * sb / P.flop_rate + 3 * P.unit_overhead);
    Tk[t].unit_cost = chain / std::max(1, Tk[t].nunits);
}

void lx_planner::factor_rl(int s)
{
    const sysinfo &sy = S[s];
    const int n = sy.n, b = sy.b, mt = sy.mt;
    const int tail = P.wide_tail < 0 ? mt / 2 : P.wide_tail;
    // Grouped updates (update_group = g > 1): the steps come in groups
    // k0 .. k0 + g - 1 (k0 a multiple of g, the whole group before the narrow
    // tail).  Tiles with j >= k0 + g + 1 take no unit pass at the group's first
    // g - 1 steps (their GEMM(i,j,k) / SYRK(i,k) tasks keep their place in the
    // task graph with no units of their own); at the group's last step the
    // units of GEMM(i,j,k0+g-1) / SYRK(i,k0+g-1) apply all g panels at once
    // (K = g b over the contiguous panel columns), so such a tile is read and
    // written once per g updates.  No task reads those tiles during the group,
    // and every panel of the group precedes the last one's inputs
    // (TRSM(i,k) -> GEMM(i,k+1,k) -> TRSM(i,k+1), with column k + 1 always
    // updated one panel at a time).  Columns up to k0 + g keep one pass per
    // panel, so the diagonal chain's next tiles are never delayed.
    const int g = std::max(1, P.update_group);
    auto group_ok = [&](int k0) { return g > 1 && k0 % g == 0 && k0 + g - 1 < mt && mt - (k0 + g - 1) > tail; };
    auto deferred = [&](int k, int j) {
        const int k0 = k - k % g;
        return k - k0 < g - 1 && group_ok(k0) && j >= k0 + g + 1;
    };
    auto combined = [&](int k, int j) {
        const int k0 = k - k % g;
        return k - k0 == g - 1 && group_ok(k0) && j >= k0 + g + 1;
    };
    for (int k = 0; k < mt; ++k) {
        const int bs = std::min(b, n - k * b);
        // a combined unit reaches back to its group's first panel: K and the
        // operands' first column
        const int kb_pair = (k - (g - 1)) * b, bs_pair = (g - 1) * b + bs;
        {   // POTRF(k)  chol_rlmain.c:30
            int outc = tile_ctr(s, k, k);
            int t = new_task(LX_T_POTRF, s, k, k, k, PRIO_HI, outc, nullptr, 0);
            potrf_units(t, s, k * b, bs, outc);
            issued[outc] += 1;  // one unit (the last) bumps the tile's counter
            last_writer[outc] = t;
        }
        for (int i = k + 1; i < mt; ++i) {  // TRSM(i,k)  chol_rlmain.c:35-38
            const int ibs = std::min(b, n - i * b);
            int outc = tile_ctr(s, i, k), in[1] = {tile_ctr(s, k, k)};
            int t = new_task(LX_T_TRSM, s, i, k, k, PRIO_HI, outc, in, 1);
            // the lookahead tile's panel solve feeds the next diagonal factor:
            // short units; elsewhere taller ones (fewer publications)
            const int tr = i == k + 1 ? P.trsm_rows : std::max(P.trsm_rows, P.trsm_rows_wide);
            for (int r0 = 0; r0 < ibs; r0 += tr) {
                lx_unit u = proto_unit(LX_OP_TRSM_RLTN, s);
                u.M = std::min(tr, ibs - r0);
                u.N = bs;
                u.ctr = outc;
                u.c = rc(i * b + r0, k * b);
                u.a = rc(k * b, k * b);
                add_unit(t, u);
            }
            Tk[t].unit_cost = (double)std::min(tr, ibs) * bs * bs / P.trsm_rate + P.unit_overhead;
            issued[outc] += Tk[t].nunits;
            last_writer[outc] = t;
        }
        for (int i = k + 1; i < mt; ++i) {
            const int ibs = std::min(b, n - i * b);
            for (int j = k + 1; j < i; ++j) {  // GEMM(i,j,k)  chol_rlmain.c:43-46
                const int jbs = std::min(b, n - j * b);
                int outc = tile_ctr(s, i, j), in[2] = {tile_ctr(s, i, k), tile_ctr(s, j, k)};
                int t = new_task(LX_T_GEMM, s, i, j, k, PRIO_LO, outc, in, 2);
                if (deferred(k, j)) {  // applied by the group's last GEMM(i,j,.) units
                    last_writer[outc] = t;
                    continue;
                }
                const bool comb = combined(k, j);
                // the next panel's column (j = k+1) feeds the critical path: fine units;
                // elsewhere wide units, so that one release fence covers more work
                // (and in the last wide_tail panels, whose trailing matrix no longer
                // gives every workgroup a wide unit)
                const bool fine = j == k + 1 || mt - k <= tail;
                const int un = fine ? P.unit_n : std::max(P.unit_n, P.gemm_wide_n);
                // each unit's publication writes its XCD's L2 back (measured at
                // about 100 us of unit time when units are 128 rows), so units
                // are as tall as the step's parallelism allows: the tallest
                // height (gemm_wide_m, halved down to unit_m) that still leaves
                // wide_units units per workgroup among the step's trailing tiles
                int um = P.unit_m;
                if (!fine) {
                    const double tiles = 0.5 * (mt - k - 1) * (mt - k - 2);
                    um = std::max(P.unit_m, P.gemm_wide_m);
                    while (um > P.unit_m &&
                           tiles * ((b + um - 1) / um) * ((b + un - 1) / un) < P.wide_units * P.workers)
                        um /= 2;
                    um = std::max(um, P.unit_m) / P.unit_m * P.unit_m;
                }
                for (int c0 = 0; c0 < jbs; c0 += un)
                    for (int r0 = 0; r0 < ibs; r0 += um) {
                        lx_unit u = proto_unit(LX_OP_GEMM_NT, s);
                        u.aux0 = upd_flag();
                        u.M = std::min(um, ibs - r0);
                        u.N = std::min(un, jbs - c0);
                        u.K = comb ? bs_pair : bs;
                        u.ctr = outc;
                        u.c = rc(i * b + r0, j * b + c0);
                        u.a = rc(i * b + r0, comb ? kb_pair : k * b);
                        u.b = rc(j * b + c0, comb ? kb_pair : k * b);
                        add_unit(t, u);
                    }
                Tk[t].unit_cost = 2.0 * um * un * (comb ? bs_pair : bs) / P.flop_rate + P.unit_overhead;
                issued[outc] += Tk[t].nunits;
                last_writer[outc] = t;
            }
            {   // SYRK(i,k)  chol_rlmain.c:48
                int outc = tile_ctr(s, i, i), in[1] = {tile_ctr(s, i, k)};
                int t = new_task(LX_T_SYRK, s, i, i, k, PRIO_LO, outc, in, 1);
                if (deferred(k, i)) {  // applied by the group's last SYRK(i,.) units
                    last_writer[outc] = t;
                    continue;
                }
                const bool comb = combined(k, i);
                for (int c0 = 0; c0 < ibs; c0 += P.unit_n)
                    for (int r0 = 0; r0 < ibs; r0 += P.unit_m) {
                        if (r0 + P.unit_m - 1 < c0) continue;  // strictly upper
                        lx_unit u = proto_unit(LX_OP_SYRK, s);
                        u.aux0 = upd_flag();
                        u.M = std::min(P.unit_m, ibs - r0);
                        u.N = std::min(P.unit_n, ibs - c0);
                        u.K = comb ? bs_pair : bs;
                        u.ctr = outc;
                        u.c = rc(i * b + r0, i * b + c0);
                        u.a = rc(i * b + r0, comb ? kb_pair : k * b);
                        u.b = rc(i * b + c0, comb ? kb_pair : k * b);
                        u.aux0 |= 1;
                        u.aux1 = r0 - c0;
                        add_unit(t, u);
                    }
                Tk[t].unit_cost = 2.0 * P.unit_m * P.unit_n * (comb ? bs_pair : bs) / P.flop_rate + P.unit_overhead;
                issued[outc] += Tk[t].nunits;
                last_writer[outc] = t;
            }
        }
    }
}

void lx_planner::factor_ll(int s)
{
    const sysinfo &sy = S[s];
    const int n = sy.n, b = sy.b, mt = sy.mt;
    for (int k = 0; k < mt; ++k) {
        const int bs = std::min(b, n - k * b);
        for (int j = 0; j < k; ++j) {  // SYRK(k; j)  chol_llmain.c:34-36
            int outc = tile_ctr(s, k, k), in[1] = {tile_ctr(s, k, j)};
            int t = new_task(LX_T_SYRK, s, k, k, j, PRIO_HI, outc, in, 1);
            for (int c0 = 0; c0 < bs; c0 += P.unit_n)
                for (int r0 = 0; r0 < bs; r0 += P.unit_m) {
                    if (r0 + P.unit_m - 1 < c0) continue;
                    lx_unit u = proto_unit(LX_OP_SYRK, s);
                    u.aux0 = upd_flag();
                    u.M = std::min(P.unit_m, bs - r0);
                    u.N = std::min(P.unit_n, bs - c0);
                    u.K = b;
                    u.ctr = outc;
                    u.c = rc(k * b + r0, k * b + c0);
                    u.a = rc(k * b + r0, j * b);
                    u.b = rc(k * b + c0, j * b);
                    u.aux0 |= 1;
                    u.aux1 = r0 - c0;
                    add_unit(t, u);
                }
            Tk[t].unit_cost = 2.0 * P.unit_m * P.unit_n * b / P.flop_rate + P.unit_overhead;
            issued[outc] += Tk[t].nunits;
            last_writer[outc] = t;
        }
        {   // POTRF(k)  chol_llmain.c:38
            int outc = tile_ctr(s, k, k);
            int t = new_task(LX_T_POTRF, s, k, k, k, PRIO_HI, outc, nullptr, 0);
            potrf_units(t, s, k * b, bs, outc);
            issued[outc] += 1;  // one unit (the last) bumps the tile's counter
            last_writer[outc] = t;
        }
        for (int i = k + 1; i < mt; ++i) {
            const int ibs = std::min(b, n - i * b);
            for (int j = 0; j < k; ++j) {  // GEMM(i,k; j)  chol_llmain.c:46-48
                int outc = tile_ctr(s, i, k), in[2] = {tile_ctr(s, i, j), tile_ctr(s, k, j)};
                int t = new_task(LX_T_GEMM, s, i, k, j, PRIO_LO, outc, in, 2);
                for (int c0 = 0; c0 < bs; c0 += P.unit_n)
                    for (int r0 = 0; r0 < ibs; r0 += P.unit_m) {
                        lx_unit u = proto_unit(LX_OP_GEMM_NT, s);
                        u.aux0 = upd_flag();
                        u.M = std::min(P.unit_m, ibs - r0);
                        u.N = std::min(P.unit_n, bs - c0);
                        u.K = b;
                        u.ctr = outc;
                        u.c = rc(i * b + r0, k * b + c0);
                        u.a = rc(i * b + r0, j * b);
                        u.b = rc(k * b + c0, j * b);
                        add_unit(t, u);
                    }
                Tk[t].unit_cost = 2.0 * P.unit_m * P.unit_n * b / P.flop_rate + P.unit_overhead;
                issued[outc] += Tk[t].nunits;
                last_writer[outc] = t;
            }
            {   // TRSM(i,k)  chol_llmain.c:50
                int outc = tile_ctr(s, i, k), in[1] = {tile_ctr(s, k, k)};
                int t = new_task(LX_T_TRSM, s, i, k, k, PRIO_LO, outc, in, 1);
                for (int r0 = 0; r0 < ibs; r0 += P.trsm_rows) {
                    lx_unit u = proto_unit(LX_OP_TRSM_RLTN, s);
                    u.M = std::min(P.trsm_rows, ibs - r0);
                    u.N = bs;
                    u.ctr = outc;
                    u.c = rc(i * b + r0, k * b);
                    u.a = rc(k * b, k * b);
                    add_unit(t, u);
                }
                Tk[t].unit_cost = (double)std::min(P.trsm_rows, ibs) * bs * bs / P.trsm_rate + P.unit_overhead;
                issued[outc] += Tk[t].nunits;
                last_writer[outc] = t;
            }
        }
    }
}

// Forward sweep, bblas_trsm.h:119-138: TRSM on block 0, then for each row
// block i the GEMMs with L(i, j<i) and the TRSM with L(i,i).
void lx_planner::sweep_forward(int s)
{
    const sysinfo &sy = S[s];
    const int n = sy.n, b = sy.b, mt = sy.mt, nrhs = sy.nrhs;
    auto trsv = [&](int i) {
        const int ibs = std::min(b, n - i * b);
        int outc = rhs_ctr(s, i), in[1] = {tile_ctr(s, i, i)};
        int t = new_task(LX_T_SV_TRSM, s, i, i, 0, PRIO_SWEEP, outc, in, 1);
        lx_unit u = proto_unit(LX_OP_TRSV, s);
        u.M = ibs;
        u.N = nrhs;
        u.ctr = outc;
        u.cbase = LX_BASE_X;
        u.c = rc(i * b, 0);
        u.a = rc(i * b, i * b);
        u.aux0 = 0;
        add_unit(t, u);
        Tk[t].unit_cost = 2.0 * ibs * ibs * P.elem_bytes / P.byte_rate + P.unit_overhead;
        issued[outc] += 1;
        last_writer[outc] = t;
    };
    trsv(0);
    for (int i = 1; i < mt; ++i) {
        const int ibs = std::min(b, n - i * b);
        for (int j = 0; j < i; ++j) {
            const int jbs = std::min(b, n - j * b);
            int outc = rhs_ctr(s, i), in[2] = {tile_ctr(s, i, j), rhs_ctr(s, j)};
            int t = new_task(LX_T_SV_GEMM, s, i, j, 0, PRIO_SWEEP, outc, in, 2);
            for (int r0 = 0; r0 < ibs; r0 += P.gemv_rows) {
                lx_unit u = proto_unit(LX_OP_GEMV_RHS, s);
                u.M = std::min(P.gemv_rows, ibs - r0);
                u.K = jbs;
                u.N = nrhs;
                u.ctr = outc;
                u.cbase = LX_BASE_X;
                u.bbase = LX_BASE_X;
                u.c = rc(i * b + r0, 0);
                u.a = rc(i * b + r0, j * b);
                u.b = rc(j * b, 0);
                u.aux0 = 0;
                add_unit(t, u);
            }
            Tk[t].unit_cost = (double)std::min(P.gemv_rows, ibs) * jbs * P.elem_bytes / P.byte_rate + P.unit_overhead;
            issued[outc] += Tk[t].nunits;
            last_writer[outc] = t;
        }
        trsv(i);
    }
}

// Backward sweep, bblas_trsm.h:139-160: TRSM on the (possibly ragged) last
// block, then walking up: GEMMs with L(last..j+1, j)^T, then TRSM with L(j,j)^T.
void lx_planner::sweep_backward(int s)
{
    const sysinfo &sy = S[s];
    const int n = sy.n, b = sy.b, mt = sy.mt, nrhs = sy.nrhs;
    auto trsv = [&](int i) {
        const int ibs = std::min(b, n - i * b);
        int outc = rhs_ctr(s, i), in[1] = {tile_ctr(s, i, i)};
        int t = new_task(LX_T_SV_TRSM, s, i, i, 1, PRIO_SWEEP, outc, in, 1);
        lx_unit u = proto_unit(LX_OP_TRSV, s);
        u.M = ibs;
        u.N = nrhs;
        u.ctr = outc;
        u.cbase = LX_BASE_X;
        u.c = rc(i * b, 0);
        u.a = rc(i * b, i * b);
        u.aux0 = 1;
        add_unit(t, u);
        Tk[t].unit_cost = 2.0 * ibs * ibs * P.elem_bytes / P.byte_rate + P.unit_overhead;
        issued[outc] += 1;
        last_writer[outc] = t;
    };
    trsv(mt - 1);
    for (int j = mt - 2; j >= 0; --j) {
        const int jbs = std::min(b, n - j * b);
        for (int i = mt - 1; i > j; --i) {
            const int ibs = std::min(b, n - i * b);
            int outc = rhs_ctr(s, j), in[2] = {tile_ctr(s, i, j), rhs_ctr(s, i)};
            int t = new_task(LX_T_SV_GEMM, s, j, i, 1, PRIO_SWEEP, outc, in, 2);
            for (int r0 = 0; r0 < jbs; r0 += P.gemv_rows) {
                lx_unit u = proto_unit(LX_OP_GEMV_RHS, s);
                u.M = std::min(P.gemv_rows, jbs - r0);
                u.K = ibs;
                u.N = nrhs;
                u.ctr = outc;
                u.cbase = LX_BASE_X;
                u.bbase = LX_BASE_X;
                u.c = rc(j * b + r0, 0);
                u.a = rc(i * b, j * b + r0);
                u.b = rc(i * b, 0);
                u.aux0 = 1;
                add_unit(t, u);
            }
            Tk[t].unit_cost = (double)std::min(P.gemv_rows, jbs) * ibs * P.elem_bytes / P.byte_rate + P.unit_overhead;
            issued[outc] += Tk[t].nunits;
            last_writer[outc] = t;
        }
        trsv(j);
    }
}

void lx_planner::finalize()
{
    if (finalized) return;
    finalized = true;
    const int nt = (int)Tk.size();
    std::vector<std::vector<int32_t>> succ(nt);
    std::vector<int32_t> pending(nt, 0);
    for (int t = 0; t < nt; ++t) {
        pending[t] = (int)Tk[t].preds.size();
        for (int p : Tk[t].preds) succ[p].push_back(t);
    }
    // Bottom levels (tasks are created in a topological order).
    std::vector<double> bl(nt, 0.0);
    for (int t = nt - 1; t >= 0; --t) {
        double m = 0.0;
        for (int q : succ[t]) m = std::max(m, bl[q]);
        bl[t] = Tk[t].unit_cost + m;
    }
    // Priority of a task in the simulated ready queue.
    auto key = [&](int t) -> double { return P.policy == 0 ? (double)Tk[t].prio : bl[t]; };
    struct RI {
        double k;
        int32_t task, unit;
    };
    auto rcmp = [](const RI &a, const RI &b) {
        if (a.k != b.k) return a.k < b.k;           // larger key first
        if (a.task != b.task) return a.task > b.task;  // earlier task first
        return a.unit > b.unit;
    };
    std::priority_queue<RI, std::vector<RI>, decltype(rcmp)> ready(rcmp);
    struct EV {
        double t;
        int32_t unit;
    };
    auto ecmp = [](const EV &a, const EV &b) { return a.t > b.t; };
    std::priority_queue<EV, std::vector<EV>, decltype(ecmp)> events(ecmp);
    // a task with no units of its own (its work done by a later task's units)
    // completes as soon as it is ready
    std::function<void(int)> push_task = [&](int t) {
        if (Tk[t].nunits == 0) {
            for (int q : succ[t])
                if (--pending[q] == 0) push_task(q);
            return;
        }
        for (int q = 0; q < Tk[t].nunits; ++q) ready.push({key(t), t, Tk[t].first_unit + q});
    };
    for (int t = 0; t < nt; ++t)
        if (pending[t] == 0) push_task(t);
    std::vector<int32_t> done_units(nt, 0);
    std::vector<int32_t> order;
    order.reserve(U.size());
    double now = 0.0;
    int free_w = std::max(1, P.workers);
    while (!ready.empty() || !events.empty()) {
        while (free_w > 0 && !ready.empty()) {
            RI r = ready.top();
            ready.pop();
            order.push_back(r.unit);
            events.push({now + Tk[r.task].unit_cost, r.unit});
            --free_w;
        }
        if (events.empty()) break;
        EV e = events.top();
        events.pop();
        now = e.t;
        ++free_w;
        int t = unit_task[e.unit];
        if (++done_units[t] == Tk[t].nunits)
            for (int q : succ[t])
                if (--pending[q] == 0) push_task(q);
    }
    if (order.size() != U.size()) throw std::logic_error("plan simulation did not schedule every unit");
    sim_time = now;
    std::vector<lx_unit> ordered;
    ordered.reserve(U.size());
    for (int32_t q : order) ordered.push_back(U[q]);
    U.swap(ordered);
}

// True when each unit's operands (C, and A / B where the op reads them) lie
// inside single b x b tiles: the condition for running the plan on a
// hypermatrix, whose tiles are separate buffers.
bool lx_units_in_tiles(const std::vector<lx_unit> &U, int b)
{
    auto inside = [&](int64_t rcp, int rows, int cols) {
        const int64_t r = rcp & 0xffffffffll, c = rcp >> 32;
        return rows > 0 && cols > 0 && r / b == (r + rows - 1) / b && c / b == (c + cols - 1) / b;
    };
    for (const lx_unit &u : U) {
        bool ok = false;
        switch (u.op) {
        case LX_OP_POTRF: ok = inside(u.c, u.M, u.M); break;
        case LX_OP_TRSM_RLTN: ok = inside(u.c, u.M, u.N) && inside(u.a, u.N, u.N); break;
        case LX_OP_SYRK:
        case LX_OP_GEMM_NT: ok = inside(u.c, u.M, u.N) && inside(u.a, u.M, u.K) && inside(u.b, u.N, u.K); break;
        default: break;
        }
        if (!ok) return false;
    }
    return true;
}
