"""GPU parity of the SPD solve, the blocked sweeps and the refinement loop.

Reference: dspdsolv_main (src/spdsolv/spdsolv_main.c:10-20), bblas_dltrsm
lower branches (src/blocked_blas/bblas_trsm.h:116-160), itref's loop
(src/itref/itref.c:999-1101).  Bars: FP64 normwise backward error
||b - A x||_inf / (||A||_inf ||x||_inf) <= 1e-14; sweeps against the oracle
within the trsm_test threshold style (relative 1-norm error).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _berr(A, X, B):
    A = np.asarray(A)
    R = B - A @ X
    an = np.abs(A).sum(axis=1).max()
    return max(np.abs(R[:, c]).max() / (an * np.abs(X[:, c]).max()) for c in range(X.shape[1]))


def _spd_gauss(n, seed, alpha):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return np.asfortranarray(G @ G.T / n + alpha * np.eye(n))


@pytest.mark.parametrize("m,nr,b", [(8, 1, 4), (130, 1, 32), (512, 3, 128), (1989, 2, 533), (4096, 1, 256)])
def test_dspdsolv(lx, orc, dev, m, nr, b):
    print(f"[spdsolv] m={m} nr={nr} b={b}", flush=True)
    A0 = orc.genmat_spd(m, seed=0, diag_mode=1)
    X0 = np.ones((m, nr), order="F")
    B0 = np.asfortranarray(A0 @ X0)
    A, B = A0.copy(order="F"), B0.copy(order="F")
    assert lx.dspdsolv_main(A, B, b) == 0
    assert _berr(A0, B, B0) <= 1e-14
    # same result as the oracle's restatement to FP64 accuracy
    Ao, Bo = A0.copy(order="F"), B0.copy(order="F")
    assert orc.dspdsolv(Ao, Bo, b) == 0
    assert np.abs(B - Bo).max() / np.abs(Bo).max() <= 1e-13
    assert np.linalg.norm(np.tril(A) - np.tril(Ao)) / np.linalg.norm(np.tril(Ao)) <= 1e-12


@pytest.mark.parametrize("m,b", [(700, 128), (1024, 256)])
def test_sspdsolv(lx, orc, dev, m, b):
    print(f"[sspdsolv] m={m} b={b}", flush=True)
    A0 = orc.genmat_spd(m, seed=1, diag_mode=1)
    B0 = np.asfortranarray(A0 @ np.ones((m, 2)))
    A, B = A0.astype(np.float32, order="F"), B0.astype(np.float32, order="F")
    assert lx.sspdsolv_main(A, B, b) == 0
    assert _berr(A0, B.astype(np.float64), B0) <= 1e-6


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,nr,bm", [(600, 5, 121), (500, 1, 128), (64, 3, 64), (1000, 16, 256)])
def test_dltrsm(lx, orc, dev, trans, m, nr, bm):
    """trsm_test.c's ragged shapes (600 x 500 with 121 tiles) for the lower branches."""
    print(f"[ltrsm {trans}] m={m} nr={nr} bm={bm}", flush=True)
    rng = np.random.default_rng(4)
    L = np.asfortranarray(np.tril(rng.uniform(0, 1, (m, m))) + m * np.eye(m))
    B0 = np.asfortranarray(rng.uniform(-1, 1, (m, nr)))
    alpha = -2.11
    B = B0.copy(order="F")
    lx.dltrsm(trans, L, B, bm, alpha)
    Bo = B0.copy(order="F")
    orc.ltrsm_lower(0 if trans == "N" else 1, L, Bo, bm, alpha)
    rel = np.abs(B - Bo).sum(axis=0).max() / np.abs(Bo).sum(axis=0).max()
    assert rel <= 1e-14, rel


@pytest.mark.parametrize("n,nr,b,alpha", [(512, 1, 128, 1e-2), (1000, 4, 256, 1e-3), (2048, 16, 512, 1e-4)])
def test_dposv_ir(lx, orc, dev, n, nr, b, alpha):
    print(f"[ir] n={n} nr={nr} b={b} alpha={alpha}", flush=True)
    A = _spd_gauss(n, 1, alpha)
    rng = np.random.default_rng(2)
    B = np.asfortranarray(rng.uniform(-1, 1, (n, nr)))
    X, st = lx.dposv_ir(A, B, b)
    assert st.info == 0
    assert st.fallback == 0
    assert st.berr <= 1e-14
    assert st.sweeps <= 4  # an FP32 factor at cond ~1e4 needs about two corrections
    assert _berr(A, X, B) <= 1e-14
    Xo, sto, info = orc.dposv_ir(A, B, b)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() <= 1e-10


@pytest.mark.parametrize("update_prec", [0, 2])
def test_dposv_ir_zero_column(lx, orc, dev, update_prec):
    """A zero right-hand side among several: x = 0 and r = 0 there, which is
    an exact solution (backward error 0, not 0 / 0), so the loop converges on
    the other columns without the FP64 fallback."""
    n, b = 800, 128
    print(f"[ir zero column] update_prec={update_prec}", flush=True)
    A = _spd_gauss(n, 6, 1e-2)
    B = np.asfortranarray(np.random.default_rng(7).uniform(-1, 1, (n, 4)))
    B[:, 2] = 0.0
    X, st = lx.dposv_ir(A, B, b, lx.IrOpts(30, 1e-14, 0.5, update_prec))
    assert st.info == 0 and st.fallback == 0
    assert np.isfinite(st.berr) and st.berr <= 1e-14
    assert not X[:, 2].any()
    assert _berr(A, X[:, [0, 1, 3]], B[:, [0, 1, 3]]) <= 1e-14


def test_dposv_ir_many_rhs(lx, orc, dev):
    """More than 16 right-hand sides (the sweeps take 16 at a time)."""
    n, nr, b = 1100, 37, 256
    print(f"[ir many rhs] n={n} nr={nr}", flush=True)
    A = _spd_gauss(n, 8, 1e-3)
    B = np.asfortranarray(np.random.default_rng(9).uniform(-1, 1, (n, nr)))
    X, st = lx.dposv_ir(A, B, b, lx.IrOpts(30, 1e-14, 0.5, 2))
    assert st.info == 0 and st.fallback == 0 and st.berr <= 1e-14
    assert _berr(A, X, B) <= 1e-14
    Xo, sto, info = orc.dposv_ir(A, B, b)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() <= 2.5 * (4 + 1e-3) / 1e-3 * 1e-14


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("nr", [8, 16])
def test_sltrsm_multi_rhs(lx, orc, dev, trans, nr):
    """FP32 sweeps with several right-hand sides (partial sums on the matrix
    cores) against the oracle's blocked sweeps in FP64 on the same FP32
    inputs, on a ragged order (1000 = 3 x 256 + 232)."""
    m, bm = 1000, 256
    print(f"[sltrsm {trans}] m={m} nr={nr}", flush=True)
    rng = np.random.default_rng(11)
    L = np.asfortranarray(np.tril(rng.uniform(0, 1, (m, m))) + 0.1 * m * np.eye(m)).astype(np.float32)
    B0 = np.asfortranarray(rng.uniform(-1, 1, (m, nr))).astype(np.float32)
    B = B0.copy(order="F")
    lx.sltrsm(trans, L, B, bm, 1.5)
    Bo = B0.astype(np.float64, order="F")
    orc.ltrsm_lower(0 if trans == "N" else 1, np.asfortranarray(L.astype(np.float64)), Bo, bm, 1.5)
    rel = np.abs(B - Bo).sum(axis=0).max() / np.abs(Bo).sum(axis=0).max()
    assert rel <= 2e-5, rel


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,nr", [(300, 1), (1000, 1), (4100, 1), (8192, 1), (1000, 5), (4100, 16), (2000, 21)])
def test_sltrsm_128_row_steps(lx, orc, dev, monkeypatch, trans, m, nr):
    """FP32 sweeps with 128-row steps (the step matrices' pre-pass, then the
    one-column kernel or, for several right-hand sides, the 16-column kernel
    on the matrix cores; forced down to any order here) against the oracle's
    blocked sweeps in FP64 on the same FP32 inputs, on a Cholesky factor of a
    Gram matrix plus 1e-2 I; ragged orders and a second 16-column launch
    (21 columns) included.  The 64-row kernel on the same inputs sets the
    scale: the error may not exceed twice its error or the FP32 bound."""
    print(f"[sltrsm 128 {trans}] m={m} nr={nr}", flush=True)
    rng = np.random.default_rng(m)
    G = rng.standard_normal((m, m))
    L = np.asfortranarray(np.linalg.cholesky(G @ G.T / m + 1e-2 * np.eye(m))).astype(np.float32)
    B0 = np.asfortranarray(rng.uniform(-1, 1, (m, nr))).astype(np.float32)
    Bo = B0.astype(np.float64, order="F")
    orc.ltrsm_lower(0 if trans == "N" else 1, np.asfortranarray(L.astype(np.float64)), Bo, 256, 1.0)
    errs = {}
    for min_n in ("1", "0"):  # 128-row steps, then the 64-row kernel
        monkeypatch.setenv("LINOMPSS_SWEEP2_MIN_N", min_n)
        B = B0.copy(order="F")
        lx.sltrsm(trans, L, B, 256, 1.0)
        assert np.isfinite(B).all()
        errs[min_n] = np.abs(B - Bo).max() / np.abs(Bo).max()
    print(f"  rel err 128-row {errs['1']:.2e}, 64-row {errs['0']:.2e}", flush=True)
    assert errs["1"] <= max(2 * errs["0"], 2e-6), errs


def test_dposv_ir_fallback(lx, dev):
    """cond(A) ~ 1e9 is beyond what an FP32 factor can refine: the FP64 factor answers."""
    n, b = 600, 128
    print("[ir fallback]", flush=True)
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, -9, n)
    A = np.asfortranarray((Q * ev) @ Q.T)
    A = np.asfortranarray((A + A.T) / 2)
    B = np.asfortranarray(rng.uniform(-1, 1, (n, 1)))
    X, st = lx.dposv_ir(A, B, b)
    assert st.fallback == 1
    assert _berr(A, X, B) <= 1e-14


def test_device_posv_single_launch(lx, orc, dev):
    n, nr, b = 2048, 2, 256
    print("[posv dev]", flush=True)
    A0 = orc.genmat_spd(n, seed=6, diag_mode=1)
    B0 = np.asfortranarray(A0 @ np.ones((n, nr)))
    dA = lx.DeviceMatrix.from_host(A0)
    dX = lx.DeviceMatrix.from_host(B0)
    assert lx.posv_dev(dA, dX, b) == 0
    X = dX.to_host()
    assert _berr(A0, X, B0) <= 1e-14


@pytest.mark.parametrize("n,nr,b,alpha", [(1000, 1, 256, 1e-3), (2048, 4, 512, 1e-4)])
def test_dposv_ir_split_bf16_updates(lx, orc, dev, n, nr, b, alpha):
    """Refinement with the FP32 factor's trailing SYRK/GEMM on split-BF16
    matrix cores (update_prec = 1): same bar as the FP32 factor."""
    print(f"[ir bf16x6] n={n} nr={nr} b={b}", flush=True)
    A = _spd_gauss(n, 3, alpha)
    rng = np.random.default_rng(4)
    B = np.asfortranarray(rng.uniform(-1, 1, (n, nr)))
    X, st = lx.dposv_ir(A, B, b, lx.IrOpts(30, 1e-14, 0.5, 1))
    assert st.info == 0
    assert st.fallback == 0
    assert st.berr <= 1e-14
    assert _berr(A, X, B) <= 1e-14
    # two solutions with backward error <= 1e-14 differ by at most ~2 cond(A) 1e-14
    # (cond ~ (4 + alpha) / alpha here)
    Xo, sto, info = orc.dposv_ir(A, B, b)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() <= 2.5 * (4 + alpha) / alpha * 1e-14


@pytest.mark.parametrize("n,nr,b,alpha,scale", [(1000, 1, 256, 1e-3, 0), (2048, 4, 512, 1e-4, 0),
                                                (1500, 2, 256, 1e-3, 12), (1200, 2, 256, 1e-3, 150)])
def test_dposv_ir_split_fp16_updates(lx, orc, dev, n, nr, b, alpha, scale):
    """Refinement with the trailing SYRK/GEMM on split-FP16 matrix cores
    (update_prec = 2): the driver factors D A D (powers of two, diagonal in
    [1/4, 1)), so a badly scaled A (rows scaled by up to 2^scale) converges
    like a well scaled one, without the FP64 fallback.  At 2^150 the FP32
    right-hand sides must be scaled by max |D r|, not max |r| (D reaches 2^150,
    past FP32's range)."""
    print(f"[ir fp16x3] n={n} nr={nr} b={b} scale=2^{scale}", flush=True)
    A = _spd_gauss(n, 3, alpha)
    if scale:
        S = 2.0 ** np.random.default_rng(5).integers(-scale, scale + 1, n)
        A = np.asfortranarray(S[:, None] * A * S[None, :])
    rng = np.random.default_rng(4)
    B = np.asfortranarray(rng.uniform(-1, 1, (n, nr)))
    X, st = lx.dposv_ir(A, B, b, lx.IrOpts(30, 1e-14, 0.5, 2))
    assert st.info == 0
    assert st.fallback == 0
    assert st.berr <= 1e-14
    assert _berr(A, X, B) <= 1e-14
    if not scale:
        Xo, sto, info = orc.dposv_ir(A, B, b)
        assert np.abs(X - Xo).max() / np.abs(Xo).max() <= 2.5 * (4 + alpha) / alpha * 1e-14


def test_dposv_ir_from_files(lx, dev, tmp_path):
    """The refinement app's file path (src/itref/itref_setup.c:98-121,
    155-199): A.dat and RHS.dat in the Octave text format drive the solve.
    The FP64 values read back exactly, so the solve sees the same system."""
    n, s, b = 700, 3, 128
    print(f"[itref files] n={n} s={s} b={b}", flush=True)
    A0 = _spd_gauss(n, 11, 1.0)
    B0 = np.asfortranarray(np.random.default_rng(12).uniform(-1, 1, (n, s)))
    lx.write_dense(tmp_path / "A.dat", A0, "A")
    lx.write_dense(tmp_path / "RHS.dat", B0, "RHS")
    A = lx.read_dense(tmp_path / "A.dat")
    B = lx.read_dense(tmp_path / "RHS.dat")
    assert np.array_equal(A, A0) and np.array_equal(B, B0)
    X, st = lx.dposv_ir(A, B, b)
    assert _berr(A0, X, B0) <= 1e-14
