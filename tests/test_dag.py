"""The planner's task graph against the reference's recorded DAG.

visualize/TDG/dpotrf_tdg.dot is the task graph Nanos++ recorded for the
reference's right-looking dpotrf (chol_rlmain.c:25-50) at 8 x 8 tiles; its
node ids follow spawn order.  The planner walks the same loop nest, so task t
of the plan must be node t + 2 (ids 0/1 are the runtime's own), with the same
type and the same dependency edges.  The in-degree closed forms cover larger
tile counts.
"""
import json
from collections import Counter
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "dpotrf_tdg_mt8.json"
TYPES = {0: "potrf", 1: "trsm", 2: "gemm", 3: "syrk", 4: "sv_trsm", 5: "sv_gemm"}


def test_rl_plan_matches_recorded_dag(lx):
    gold = json.loads(GOLD.read_text())
    types, coords, edges = lx.plan_tdg(0, 8 * 64, 64)
    assert len(types) == len(gold["nodes"]) == 120
    for t, ty in enumerate(types):
        assert gold["nodes"][str(t + 2)] == TYPES[int(ty)], t
    mine = sorted([int(a) + 2, int(b) + 2] for a, b in edges)
    assert mine == gold["edges"]
    assert len(mine) == 252


def test_edge_type_histogram(lx):
    types, coords, edges = lx.plan_tdg(0, 8 * 32, 32)
    h = Counter((TYPES[int(types[a])], TYPES[int(types[b])]) for a, b in edges)
    assert h == {("potrf", "trsm"): 28, ("gemm", "trsm"): 21, ("trsm", "gemm"): 112, ("gemm", "gemm"): 35,
                 ("trsm", "syrk"): 28, ("syrk", "syrk"): 21, ("syrk", "potrf"): 7}


@pytest.mark.parametrize("mt", [1, 2, 5, 16, 33])
def test_rl_indegree_closed_form(lx, mt):
    """POTRF(k)=[k>0]; TRSM(i,k)=1+[k>0]; SYRK(i;k)=1+[k>0]; GEMM(i,j;k)=2+[k>0]."""
    b = 16
    types, coords, edges = lx.plan_tdg(0, mt * b - (b // 2 if mt > 1 else 0), b)
    indeg = np.zeros(len(types), dtype=int)
    for a, c in edges:
        indeg[c] += 1
    for t, ty in enumerate(types):
        k = coords[t, 2]
        expect = {0: int(k > 0), 1: 1 + int(k > 0), 3: 1 + int(k > 0), 2: 2 + int(k > 0)}[int(ty)]
        assert indeg[t] == expect, (t, ty, coords[t])
    n_t = Counter(int(x) for x in types)
    assert n_t[0] == mt and n_t[1] == mt * (mt - 1) // 2 and n_t[3] == mt * (mt - 1) // 2
    assert n_t[2] == mt * (mt - 1) * (mt - 2) // 6


def test_ll_plan_same_tasks_as_rl(lx):
    """CHOL_LL (chol_llmain.c:26-53) runs the same set of tile tasks in another order."""
    for mt in (3, 8):
        t_rl, c_rl, _ = lx.plan_tdg(0, mt * 16, 16)
        t_ll, c_ll, e_ll = lx.plan_tdg(1, mt * 16, 16)
        s_rl = sorted((int(a), *map(int, b)) for a, b in zip(t_rl, c_rl))
        s_ll = sorted((int(a), *map(int, b)) for a, b in zip(t_ll, c_ll))
        assert s_rl == s_ll
        # every LL edge goes from an earlier task to a later one
        assert all(a < b for a, b in e_ll)


def test_solve_plan_sweeps(lx):
    """Forward then backward sweep (bblas_trsm.h:119-160): mt TRSM and mt(mt-1)/2 GEMM tasks each."""
    mt = 6
    types, coords, edges = lx.plan_tdg(2, mt * 32 - 5, 32, 1)
    n_t = Counter(int(x) for x in types)
    assert n_t[4] == 2 * mt and n_t[5] == mt * (mt - 1)
    assert all(a < b for a, b in edges)


def test_plan_info_counts(lx):
    units, tasks, sim = lx.plan_info(0, 32768, 512)
    assert tasks == 45760  # 64 POTRF + 2016 TRSM + 2016 SYRK + 41664 GEMM
    assert units > tasks and sim > 0
