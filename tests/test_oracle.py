"""Pins the CPU oracle (the checker) before anything is checked against it.

Each restated function is checked against something it cannot share a mistake
with: glibc's drand48, a pure-Python fill loop, numpy/LAPACK (OpenBLAS in
numpy, not the MKL the oracle links), scipy's triangular solve, brute-force
triple loops, the committed golden fixtures, closed forms, and planted
mistakes that the checks must catch.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sl

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "chol_small.npz")


def test_drand48_bit_exact_vs_glibc(orc):
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 7, 123456789):
        libc.srand48(seed)
        ref = np.array([libc.drand48() for _ in range(1000)])
        assert np.array_equal(orc.drand48_fill(seed, 1000), ref)
    assert np.array_equal(orc.drand48_fill(7, 16), GOLD["drand48_seed7"])


def test_genmat_spd_fill_order(orc):
    """genmat.c:59-67: j outer, i >= j inner, mirrored; diagonal = draw + n or n."""
    n = 9
    r = orc.drand48_fill(3, n * (n + 1) // 2)
    A0 = orc.genmat_spd(n, 3, diag_mode=0)
    A1 = orc.genmat_spd(n, 3, diag_mode=1)
    q = 0
    for j in range(n):
        for i in range(j, n):
            exp0 = r[q] + n if i == j else r[q]
            assert A0[i, j] == A0[j, i] == exp0
            assert A1[i, j] == (n if i == j else r[q])
            q += 1
    G = orc.genmat(4, 3, 5)
    assert np.array_equal(G, orc.drand48_fill(5, 12).reshape(3, 4).T)
    S = orc.genmat_sym_full(5, 2)
    assert np.allclose(np.diag(S) - 50.0, orc.drand48_fill(2, 15)[[0, 5, 9, 12, 14]])


def test_golden_small_fixtures(orc):
    for n in (8, 64, 130):
        A = orc.genmat_spd(n, 0, 1)
        assert np.array_equal(A, GOLD[f"A{n}"])
        L = A.copy(order="F")
        assert orc.chol(L, 32 if n > 32 else 4, "rl") == 0
        assert np.allclose(np.tril(L), GOLD[f"L{n}"], rtol=0, atol=1e-13 * n)
        B = GOLD[f"b{n}"].reshape(-1, 1).copy(order="F")
        A2 = A.copy(order="F")
        assert orc.dspdsolv(A2, B, 16) == 0
        assert np.allclose(B[:, 0], GOLD[f"x{n}"], rtol=1e-13)


def _brute_chol(A):
    n = A.shape[0]
    L = np.zeros_like(A)
    for j in range(n):
        s = A[j, j] - sum(L[j, k] ** 2 for k in range(j))
        L[j, j] = np.sqrt(s)
        for i in range(j + 1, n):
            L[i, j] = (A[i, j] - sum(L[i, k] * L[j, k] for k in range(j))) / L[j, j]
    return L


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(1, 1), (5, 2), (13, 4), (16, 16), (17, 16), (130, 32), (533, 128), (1989, 533)])
def test_chol_vs_lapack(orc, n, b, variant):
    A0 = orc.genmat_spd(n, 1, diag_mode=0)
    A = A0.copy(order="F")
    assert orc.chol(A, b, variant) == 0
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(A) - Lref) <= 1e-13 * np.linalg.norm(Lref)
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    assert orc.potrf_resid(A0, A) < 10.0
    if n <= 17:
        assert np.allclose(np.tril(A), _brute_chol(A0), rtol=1e-12, atol=1e-14)


def test_schol_vs_lapack(orc):
    A0 = orc.genmat_spd(300, 4, diag_mode=1)
    for variant in ("rl", "ll"):
        Af = A0.astype(np.float32, order="F")
        assert orc.chol(Af, 64, variant) == 0
        Lref = np.linalg.cholesky(A0)
        assert np.linalg.norm(np.tril(Af) - Lref) / np.linalg.norm(Lref) < 1e-6


def test_chol_reports_info(orc):
    A = orc.genmat_spd(100, 2, diag_mode=1)
    A[40, 40] = -1.0
    assert orc.chol(A.copy(order="F"), 16, "rl") == 41
    assert orc.chol(A.copy(order="F"), 16, "ll") == 41


@pytest.mark.parametrize("variant", ["hrl", "hll"])
@pytest.mark.parametrize("mt,b", [(1, 1), (1, 7), (4, 3), (5, 16), (3, 100)])
def test_chol_hypermatrix_vs_lapack(orc, mt, b, variant):
    """CHOL_HRL / CHOL_HLL over separate tile buffers against LAPACK on the
    assembled matrix (brute force below n = 17), the tiles above the diagonal
    untouched by construction (they are not passed), the strict upper part of
    each diagonal tile left as it was; FP32 at FP32 accuracy."""
    n = mt * b
    A0 = orc.genmat_spd(n, 5, diag_mode=0)
    tiles = orc.to_tiles(A0, b)
    assert orc.chol_tiles(tiles, mt, b, variant) == 0
    L = np.tril(orc.from_tiles(tiles, mt, b))
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(L - Lref) <= 1e-13 * np.linalg.norm(Lref)
    if n <= 17:
        assert np.allclose(L, _brute_chol(A0), rtol=1e-12, atol=1e-14)
    for k in range(mt):
        assert np.array_equal(np.triu(tiles[(k, k)], 1), np.triu(A0[k * b:(k + 1) * b, k * b:(k + 1) * b], 1))
    tf = orc.to_tiles(A0.astype(np.float32), b)
    assert orc.chol_tiles(tf, mt, b, variant) == 0
    assert np.linalg.norm(np.tril(orc.from_tiles(tf, mt, b)) - Lref) <= 1e-6 * np.linalg.norm(Lref)


def test_chol_hypermatrix_reports_info(orc):
    A = orc.genmat_spd(100, 2, diag_mode=1)
    A[40, 40] = -1.0
    for variant in ("hrl", "hll"):
        assert orc.chol_tiles(orc.to_tiles(A, 20), 5, 20, variant) == 41


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("m,nr,bm", [(600, 5, 121), (64, 1, 64), (65, 3, 64), (1000, 7, 256)])
def test_ltrsm_vs_scipy(orc, trans, m, nr, bm):
    rng = np.random.default_rng(m + nr)
    L = np.asfortranarray(np.tril(rng.uniform(0, 1, (m, m))) + m * np.eye(m))
    B = np.asfortranarray(rng.uniform(-1, 1, (m, nr)))
    X = B.copy(order="F")
    orc.ltrsm_lower(trans, L, X, bm, -2.11)
    ref = sl.solve_triangular(L, -2.11 * B, lower=True, trans=trans)
    assert np.abs(X - ref).max() <= 1e-13 * np.abs(ref).max()


def test_spdsolv_backward_error(orc):
    m = 777
    A0 = orc.genmat_spd(m, 0, diag_mode=1)
    X0 = orc.genmat(m, 3, 9)
    B = np.asfortranarray(A0 @ X0)
    A = A0.copy(order="F")
    assert orc.dspdsolv(A, B, 100) == 0
    assert orc.backward_error(A0, B, np.asfortranarray(A0 @ X0)) <= 1e-14
    assert np.abs(B - X0).max() <= 1e-12


def test_backward_error_closed_form(orc):
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.standard_normal((50, 50)))
    X = np.asfortranarray(rng.standard_normal((50, 2)))
    B = np.asfortranarray(A @ X + 1e-3 * rng.standard_normal((50, 2)))
    an = np.abs(A).sum(axis=1).max()
    exp = max(np.abs(B[:, c] - A @ X[:, c]).max() / (an * np.abs(X[:, c]).max()) for c in range(2))
    assert orc.backward_error(A, X, B) == pytest.approx(exp, rel=1e-12)


def test_relerr_and_resid_catch_planted_mistakes(orc):
    """The checks must flag a transposed factor, a dropped term and a sign error."""
    n = 60
    A0 = orc.genmat_spd(n, 5, diag_mode=1)
    L = np.linalg.cholesky(A0)
    Lf = np.asfortranarray(L)
    assert orc.potrf_resid(A0, Lf) < 10 and orc.relerr_lower(np.asfortranarray(np.tril(L)), Lf) == 0.0
    bad_T = np.asfortranarray(L.T + np.tril(L, -1) * 0)          # upper factor in the lower slot
    assert orc.potrf_resid(A0, np.asfortranarray(np.tril(bad_T))) > 10
    bad_drop = Lf.copy(order="F")
    bad_drop[n - 1, n - 2] = 0.0                                  # dropped term
    assert orc.potrf_resid(A0, bad_drop) > 10
    bad_sign = Lf.copy(order="F")
    bad_sign[5, 3] = -bad_sign[5, 3]                              # sign error
    assert orc.potrf_resid(A0, bad_sign) > 10
    # relerr is the reference's |.|-based metric: blind to sign, sensitive to magnitude
    assert orc.relerr_lower(np.asfortranarray(np.tril(L)), bad_drop) > 1e-8


def test_dposv_ir_converges_and_falls_back(orc):
    rng = np.random.default_rng(1)
    n = 400
    G = rng.standard_normal((n, n))
    A = np.asfortranarray(G @ G.T / n + 1e-3 * np.eye(n))
    B = np.asfortranarray(rng.uniform(-1, 1, (n, 2)))
    X, st, info = orc.dposv_ir(A, B, 64)
    assert info == 0 and st.fallback == 0 and st.berr <= 1e-14 and st.sweeps >= 1
    assert orc.backward_error(A, X, B) <= 1e-14
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A2 = np.asfortranarray((Q * np.logspace(0, -10, n)) @ Q.T)
    A2 = np.asfortranarray((A2 + A2.T) / 2)
    X2, st2, info2 = orc.dposv_ir(A2, B, 64)
    assert st2.fallback == 1
    assert orc.backward_error(A2, X2, B) <= 1e-14
