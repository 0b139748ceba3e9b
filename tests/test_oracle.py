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


def test_genmat_covariant_entries(orc):
    """GENMAT_COVARIANT (genmat.c:263-302) with the itref app's p = k = 2
    (genmat_config.c:37): exact entries (integer powers, square roots of
    perfect squares), symmetric Toeplitz off the diagonal, and SPD with
    lambda_min >= min diag - pi^2/6 (the Toeplitz part's symbol
    2 sum cos(d t) / d^2 has minimum -pi^2/6 at t = pi)."""
    n = 40
    A = orc.genmat_covariant(n, 2.0, 2.0)
    assert A[0, 0] == 2.0 and A[3, 3] == 3.0 and A[8, 8] == 4.0 and A[24, 24] == 6.0
    assert A[0, 1] == A[1, 0] == 1.0 and A[0, 2] == 0.25 and A[7, 3] == 1.0 / 16 and A[2, 12] == 0.01
    assert np.array_equal(A, A.T)
    off = A - np.diag(np.diag(A))
    assert all(np.all(np.diag(off, d) == np.diag(off, d)[0]) for d in range(1, n))
    lam = np.linalg.eigvalsh(A)
    assert lam.min() >= 2.0 - np.pi ** 2 / 6 - 1e-12
    # p = 3, k = 3: cubes and cube roots of cubes
    B = orc.genmat_covariant(30, 3.0, 3.0)
    assert B[0, 2] == 0.125 and B[26, 26] == 4.0 and B[7, 7] == 3.0
    assert orc.genmat_covariant(5, 0.0, 2.0) is None and orc.genmat_covariant(5, -1.0, 2.0) is None


def test_genmat_costas_and_config_dispatch(orc):
    """dgenmat_config (genmat_config.c:34-133): each ompssmat_t kind reaches
    its generator (ompss_mat.h:5-14), with the default parameters {2, 2, 2, 2}
    when no configuration file is read; DIAGDOM and ONES return NULL in the
    reference.  GENMAT_COSTAS (genmat.c:199-220) maps each drand48 draw to +-1."""
    m, n = 12, 5
    r = orc.drand48_fill(4, m * n)
    C = orc.genmat_costas(m, n, 4)
    assert set(np.unique(C)) <= {-1.0, 1.0}
    assert np.array_equal(C.ravel(order="F") > 0, r > 0.5)
    assert np.array_equal(orc.dgenmat_config("costas", m, n, m, seed=4), C)
    assert np.array_equal(orc.dgenmat_config("covar", 9, 9, 9), orc.genmat_covariant(9, 2.0, 2.0))
    assert np.array_equal(orc.dgenmat_config("covar", 9, 9, 9, pars=[3.0, 3.0]), orc.genmat_covariant(9, 3.0, 3.0))
    assert np.array_equal(orc.dgenmat_config("spd", 7, 7, 9, seed=3)[:7], orc.genmat_spd(7, 3, diag_mode=0, lda=9)[:7])
    assert np.array_equal(orc.dgenmat_config("rand", 6, 4, 8, seed=5)[:6], orc.genmat(6, 4, 5))
    assert np.array_equal(orc.dgenmat_config("eye", 6, 6, 6), np.eye(6))
    assert not orc.dgenmat_config("zero", 6, 6, 7).any()
    for k in ("diagdom", "ones", 9):
        assert orc.dgenmat_config(k, 4, 4, 4) is None


def test_philox_known_answers(orc):
    """Philox4x32-10 against the known-answer vectors published with Random123
    (Salmon et al., SC'11): counter, key -> output."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        assert orc.philox4x32_10(ctr, key) == out


def test_philox_streams_distribution(orc):
    """The benchmark's uniform and Gaussian streams: range, moments and a
    Kolmogorov-Smirnov test against U[lo, hi) and N(0, 1); the two Gaussians
    of one Box-Muller pair are uncorrelated."""
    from scipy import stats
    U = orc.gen_uniform(1000, 200, 7, -1.0, 1.0).ravel()
    assert U.min() >= -1.0 and U.max() < 1.0
    assert abs(U.mean()) < 4 * np.sqrt(1 / 3 / U.size)
    assert stats.kstest(U, stats.uniform(-1, 2).cdf).pvalue > 1e-3
    G = orc.gen_gauss(1000, 200, 7).ravel(order="F")
    assert abs(G.mean()) < 4 / np.sqrt(G.size) and abs(G.var() - 1) < 4 * np.sqrt(2 / G.size)
    assert stats.kstest(G, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(G[0::2], G[1::2])[0, 1]) < 4 / np.sqrt(G.size / 2)
    # another seed gives another stream; the same seed the same
    assert not np.array_equal(orc.gen_gauss(10, 10, 8), orc.gen_gauss(10, 10, 7))
    assert np.array_equal(orc.gen_gauss(10, 10, 7), orc.gen_gauss(10, 10, 7))


def test_backward_error_zero_residual(orc):
    """A zero residual is an exact solution: backward error 0, also for the
    column b = 0, x = 0 (no 0 / 0), and the refinement loop stops at once."""
    A = orc.genmat_spd(50, 1, diag_mode=1)
    X = np.zeros((50, 2), order="F")
    X[3, 1] = 1.0  # A x = column 3 of A, exact in any summation order
    B = np.asfortranarray(A @ X)
    assert orc.backward_error(A, X, B) == 0.0
    B2 = np.asfortranarray(np.c_[np.zeros(50), np.random.default_rng(0).uniform(-1, 1, 50)])
    Xo, st, info = orc.dposv_ir(A, B2, 16)
    assert info == 0 and st.fallback == 0 and st.berr <= 1e-14
    assert not Xo[:, 0].any()
