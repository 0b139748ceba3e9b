"""GPU parity of the batched refinement solves (many independent SPD systems of
uneven order in one persistent launch).

Each system follows the single-system refinement contract (itref.c:999-1101):
FP32 factor, FP64 residual / update / backward-error test, FP64-factor
fallback on a stall.  Bars: every system's FP64 normwise backward error
||b - A x||_inf / (||A||_inf ||x||_inf) <= 1e-14 (recomputed here in numpy
from the downloaded x, and as reported by the kernel); x against the
oracle's refinement solve of the same system (dposv_ir with one tile, b = n)
within the forward-error bound 2 cond_inf(A) 1e-14; LAPACK-style info for systems
that are not positive definite.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _berr(A, x, b):
    r = b - A @ x
    return np.abs(r).max() / (np.abs(A).sum(axis=1).max() * np.abs(x).max())


def _check_solutions(lx, orc, batch, sizes, mats, rhs, st, n_oracle=6):
    X = batch.solutions()
    assert (st["info"] == 0).all()
    assert (st["fallback"] == 0).all(), st[st["fallback"] != 0]
    for i, n in enumerate(sizes):
        if n == 0:
            continue
        be = _berr(mats[i], X[i], rhs[i])
        assert be <= 1e-14, (i, n, be)
        assert st["berr"][i] <= 1e-14
        assert abs(st["berr"][i] - be) <= 1e-15 + 0.5 * be  # same quantity, summed in another order
    # against the oracle's refinement of the same systems
    pick = [i for i in np.argsort(-sizes)[:n_oracle] if sizes[i] > 0]
    for i in pick:
        n = int(sizes[i])
        Xo, sto, info = orc.dposv_ir(mats[i], rhs[i].reshape(-1, 1), n)
        assert info == 0 and sto.fallback == 0
        # both solutions have a normwise backward error <= 1e-14, so they differ by at
        # most 2 cond_inf(A) 1e-14 relative (cond_inf = ||A||_inf ||A^-1||_inf)
        cond = np.linalg.cond(mats[i], np.inf)
        err = np.abs(X[i] - Xo[:, 0]).max() / np.abs(Xo).max()
        assert err <= 2 * cond * 1e-14, (i, n, err, cond)


@pytest.mark.parametrize("use_graph,panel", [(1, 0), (0, 0), (1, 128), (1, 64)])
def test_batch_uneven(lx, orc, dev, use_graph, panel):
    print(f"[batch] 96 systems in [1, 512], graph={use_graph}, panel={panel}", flush=True)
    sizes, mats, rhs = lx.gen_batch_spd(92, 1, 512, seed=7)
    # edges: the maximum order, one row, one and two 64-row blocks
    for n in (512, 1, 64, 65):
        s1, m1, r1 = lx.gen_batch_spd(1, n, n, seed=100 + n)
        sizes = np.concatenate([s1, sizes])
        mats.insert(0, m1[0])
        rhs.insert(0, r1[0])
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    plan = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 1, 0, use_graph, panel))
    # a fixed width applies to every size class; 0 without tuning keeps the rule
    # (128-column panels above order 256)
    assert plan.info()["panel_mask"] == {0: 0xF0, 64: 0, 128: 0xFF}[panel]
    st = plan.run()
    _check_solutions(lx, orc, batch, sizes, mats, rhs, st)
    plan.close()
    batch.free()


def test_batch_graph_replays(lx, dev):
    """The captured graph replayed several times, with other device work between
    the replays, gives the same statuses and bit-identical solutions each time."""
    print("[batch] graph replays", flush=True)
    sizes, mats, rhs = lx.gen_batch_spd(300, 32, 256, seed=11)
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    plan = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 2, 0, 1))
    ref_st = plan.run()
    ref_x = batch.X.to_host().copy()
    other = lx.DeviceBatch(sizes[:50])
    other.upload(mats[:50], rhs[:50])
    for k in range(3):
        batch.X.upload(np.zeros_like(ref_x))          # eager work between replays
        lx.dposv_ir_batched_dev(other)                 # another launch of the same kernel
        st = plan.run()
        assert np.array_equal(st, ref_st), k
        assert np.array_equal(batch.X.to_host(), ref_x), k
    assert (ref_st["berr"][sizes > 0] <= 1e-14).all()
    plan.close()
    batch.free()
    other.free()


def test_batch_autotune(lx, dev):
    print("[batch] autotuned schedule", flush=True)
    sizes, mats, rhs = lx.gen_batch_spd(200, 16, 300, seed=12)
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    plan = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 0, -1, 1))
    info = plan.info()
    assert info["wg_per_cu"] in (1, 2) and info["order"] in (0, 1) and info["tune_ms"] > 0
    st = plan.run()
    X = batch.solutions()
    for i in range(len(sizes)):
        assert _berr(mats[i], X[i], rhs[i]) <= 1e-14
    # a second plan of the same batch takes the cached choice without measuring
    plan2 = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 0, -1, 1))
    i2 = plan2.info()
    assert (i2["wg_per_cu"], i2["order"], i2["tune_ms"], i2["panel_mask"]) == \
        (info["wg_per_cu"], info["order"], 0.0, info["panel_mask"])
    plan.close()
    plan2.close()
    batch.free()


def test_batch_autotune_panels(lx, dev):
    """Panel widths per size class: a plan with a fixed width leaves no panel
    choice in the cache, so a later per-class plan of the same batch measures
    the classes itself (tune_ms > 0) and then shares its mask."""
    print("[batch] autotuned panel widths", flush=True)
    sizes, mats, rhs = lx.gen_batch_spd(240, 200, 400, seed=13)
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    p64 = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 0, -1, 1, 64))
    assert p64.info()["panel_mask"] == 0
    p0 = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 0, -1, 1, 0))
    i0 = p0.info()
    assert i0["tune_ms"] > 0
    # classes with no systems (orders 1-192 and 449-512) keep the rule
    assert i0["panel_mask"] & 0x07 == 0 and i0["panel_mask"] & 0x80 == 0x80
    st = p0.run()
    X = batch.solutions()
    assert (st["fallback"] == 0).all()
    for i in range(len(sizes)):
        assert _berr(mats[i], X[i], rhs[i]) <= 1e-14
    p1 = lx.BatchPlan(batch, lx.BatchOpts(30, 1e-14, 0.5, 0, -1, 1, 0))
    assert p1.info()["panel_mask"] == i0["panel_mask"] and p1.info()["tune_ms"] == 0.0
    for p in (p64, p0, p1):
        p.close()
    batch.free()


def test_batch_fallback_and_info(lx, dev):
    """A system that is not positive definite reports info > 0 (FP64 factor,
    fallback 1); one too ill-conditioned for FP32 refinement is solved by the
    FP64 factor; the others are unaffected."""
    print("[batch] fallback / info", flush=True)
    rng = np.random.default_rng(3)
    sizes = np.array([100, 100, 150, 70, 0, 1], dtype=np.int32)
    mats, rhs = [], []
    for n in sizes:
        G = rng.standard_normal((n, n))
        mats.append(np.asfortranarray(G @ G.T / max(n, 1) + 0.1 * np.eye(n)))
        rhs.append(rng.uniform(-1, 1, n))
    mats[1][5, 5] = -1.0
    Q, _ = np.linalg.qr(rng.standard_normal((150, 150)))
    mats[2] = np.asfortranarray((Q * np.logspace(0, -9, 150)) @ Q.T)
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    st = lx.dposv_ir_batched_dev(batch)
    X = batch.solutions()
    assert st["info"][1] == 6 and st["fallback"][1] == 1
    assert st["fallback"][2] == 1 and st["info"][2] == 0
    assert _berr(mats[2], X[2], rhs[2]) <= 1e-14
    for i in (0, 3, 5):
        assert st["fallback"][i] == 0 and st["info"][i] == 0
        assert _berr(mats[i], X[i], rhs[i]) <= 1e-14
    assert tuple(st[4])[:3] == (0, 0, 0)
    batch.free()


def test_batch_arguments(lx, dev):
    import ctypes
    n = (ctypes.c_int * 2)(4, 513)
    p = (ctypes.c_void_p * 2)(1, 1)
    ld = (ctypes.c_int64 * 2)(4, 513)
    h = ctypes.c_void_p()
    assert lx.lib.linompss_batch_plan_create(2, n, p, ld, p, p, None, ctypes.byref(h)) == -2
    n[1] = 8
    ld[1] = 7
    assert lx.lib.linompss_batch_plan_create(2, n, p, ld, p, p, None, ctypes.byref(h)) == -4
    assert lx.lib.linompss_batch_plan_create(-1, n, p, ld, p, p, None, ctypes.byref(h)) == -1
    assert lx.lib.linompss_batch_plan_create(0, None, None, None, None, None, None, ctypes.byref(h)) == 0
    assert lx.lib.linompss_batch_run(h, None, None) == 0
    m = ctypes.c_uint(7)
    assert lx.lib.linompss_batch_plan_panels(h, ctypes.byref(m)) == 0 and m.value == 0xF0
    assert lx.lib.linompss_batch_plan_panels(h, None) == -2
    assert lx.lib.linompss_batch_plan_destroy(h) == 0
    n[1] = 8
    ld[1] = 8
    bad = lx.BatchOpts(30, 1e-14, 0.5, 1, 0, 0, 96)  # a panel width other than 0, 64, 128
    assert lx.lib.linompss_batch_plan_create(2, n, p, ld, p, p, ctypes.byref(bad), ctypes.byref(h)) == -7


def test_batch_workload_4096(lx, orc, dev):
    """The batched workload at full size: 4096 systems of order uniform in
    [32, 512], A_i = G_i G_i^T / n_i + 0.1 I, one right-hand side each."""
    print("[batch] 4096 systems in [32, 512]", flush=True)
    sizes, mats, rhs = lx.gen_batch_spd(4096, 32, 512, seed=2)
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    plan = lx.BatchPlan(batch)
    st = plan.run()
    assert (st["info"] == 0).all() and (st["fallback"] == 0).all()
    assert st["berr"].max() <= 1e-14
    X = batch.solutions()
    for i in np.random.default_rng(0).choice(len(sizes), 64, replace=False):
        assert _berr(mats[i], X[i], rhs[i]) <= 1e-14
    plan.close()
    batch.free()
