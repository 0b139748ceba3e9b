"""GPU parity of the standalone BLAS-3 entry points (linompss_{d,s}gemm,
_{d,s}syrk, _{d,s}trsm, for ompss_* in src/gemm/ompss_mm.h:23-27,
src/syrk/ompss_syrk.c:4-12, src/trsm/ompss_trsm.c:4-12).

The cases follow the reference's own tests (tests/dense_tests/gemm_test.c,
syrk_test.c, trsm_test.c): the same sizes, leading dimensions, alpha, beta and
flag combinations, drand48-style U[0,1) data, and its check -- the 1-norm
relative difference over the whole ld x n buffer (test_aux.h:100-112) against
a library BLAS (numpy here).  Bars: FP32 1e-5 as in the reference;
FP64 1e-14 (the reference's 1e-15 holds for its MKL-on-tiles result against
MKL; an independent summation order over K = 550 differs by a few 1e-15).
TRSM on the reference's matrix (entries 100 + U[0,1), far from well
conditioned) is checked by the normwise backward error instead, and on a
well-conditioned one by the same relative difference.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _relerr(D, C):
    return np.abs(D - C).sum(axis=0).max() / np.abs(C).sum(axis=0).max()


def _op(a, t):
    return a.T if t == "T" else a


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-14), (np.float32, 1e-5)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_gemm_reference_case(lx, dev, dt, tol, ta, tb):
    print(f"[gemm] {np.dtype(dt).name} {ta}{tb}", flush=True)
    M, N, K, LDA, LDB, LDC = 600, 500, 550, 650, 625, 700
    alpha, beta = 13.4, 99.1
    rng = np.random.default_rng(1)
    A = np.asfortranarray(rng.random((LDA, max(M, K))).astype(dt))
    B = np.asfortranarray(rng.random((LDB, max(N, K))).astype(dt))
    C = np.asfortranarray(rng.random((LDC, max(M, N))).astype(dt))
    D = C.copy(order="F")
    ar, ac = (K, M) if ta == "T" else (M, K)
    br, bc = (N, K) if tb == "T" else (K, N)
    opA = _op(A[:ar, :ac].astype(np.float64), ta)
    opB = _op(B[:br, :bc].astype(np.float64), tb)
    C[:M, :N] = (alpha * opA @ opB + beta * C[:M, :N].astype(np.float64)).astype(dt)
    lx.gemm(ta, tb, M, N, K, alpha, A, B, beta, D, b=117, d=103, c=89)
    assert _relerr(D.astype(np.float64), C.astype(np.float64)) <= tol


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-14), (np.float32, 1e-5)])
@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "N"), ("L", "T"), ("U", "T")])
def test_syrk_reference_case(lx, dev, dt, tol, uplo, trans):
    print(f"[syrk] {np.dtype(dt).name} {uplo}{trans}", flush=True)
    N, K, LDA, LDC = 600, 500, 650, 625
    alpha, beta = 13.4, 99.1
    rng = np.random.default_rng(2)
    A = np.asfortranarray(rng.random((LDA, max(N, K))).astype(dt))
    C = np.asfortranarray(rng.random((LDC, N)).astype(dt))
    D = C.copy(order="F")
    opA = (A[:K, :N].T if trans == "T" else A[:N, :K]).astype(np.float64)
    full = alpha * opA @ opA.T + beta * C[:N, :N].astype(np.float64)
    tri = np.tril(np.ones((N, N), dtype=bool)) if uplo == "L" else np.triu(np.ones((N, N), dtype=bool))
    C[:N, :N][tri] = full[tri].astype(dt)
    lx.syrk(uplo, trans, N, K, alpha, A, beta, D, b=123)
    assert _relerr(D.astype(np.float64), C.astype(np.float64)) <= tol
    assert np.array_equal(D[:N, :N][~tri], C[:N, :N][~tri])  # the other triangle untouched


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("L", "T"), ("U", "N"), ("U", "T")])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("M,N", [(600, 500), (600, 7), (7, 130)])
def test_trsm_well_conditioned(lx, dev, side, uplo, trans, diag, M, N):
    """More than 16 right-hand sides take the blocked matrix-core path, up to
    16 the sweeps (left 600 x 7; right 7 x 130, i.e. 7 right-hand sides)."""
    print(f"[trsm] {side}{uplo}{trans}{diag} {M}x{N}", flush=True)
    LDA, LDB = 650, 625
    alpha = -2.11
    na = M if side == "L" else N
    rng = np.random.default_rng(3)
    # off-diagonal entries U[0,1)/na, diagonal 1 + U[0,1): well conditioned
    # whether the diagonal is read or taken as one
    A = np.asfortranarray(rng.random((LDA, na)) / na)
    A[:na, :na] += np.eye(na)
    T = np.tril(A[:na, :na]) if uplo == "L" else np.triu(A[:na, :na])
    if diag == "U":
        np.fill_diagonal(T, 1.0)
    B = np.asfortranarray(rng.random((LDB, N)))
    D = B.copy(order="F")
    opT = _op(T, trans)
    if side == "L":
        X = np.linalg.solve(opT, alpha * B[:M, :N])
    else:
        X = np.linalg.solve(opT.T, (alpha * B[:M, :N]).T).T
    C = B.copy(order="F")
    C[:M, :N] = X
    lx.trsm(side, uplo, trans, diag, M, N, alpha, A, D, b=121)
    assert _relerr(D, C) <= 1e-12
    assert np.array_equal(D[M:], B[M:])  # rows past m untouched
    if diag == "U":  # the stored diagonal is not read
        A2 = A.copy(order="F")
        np.fill_diagonal(A2[:na, :na], 1e30)
        D2 = B.copy(order="F")
        lx.trsm(side, uplo, trans, diag, M, N, alpha, A2, D2, b=121)
        assert np.array_equal(D2, D)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("L", "T"), ("U", "N"), ("U", "T")])
def test_trsm_reference_case_backward_error(lx, dev, dt, uplo, trans):
    """The reference's TRSM data (A = 100 + U[0,1), B = 100 + U[0,1), alpha =
    -2.11, M = 600, N = 500): normwise backward error of op(A) X = alpha B."""
    print(f"[trsm ref] {np.dtype(dt).name} {uplo}{trans}", flush=True)
    M, N, LDA, LDC = 600, 500, 650, 625
    alpha = -2.11
    rng = np.random.default_rng(4)
    A = np.asfortranarray((rng.random((LDA, M)) + 100).astype(dt))
    B = np.asfortranarray((rng.random((LDC, N)) + 100).astype(dt))
    D = B.copy(order="F")
    lx.trsm("L", uplo, trans, "N", M, N, alpha, A, D, b=121)
    T = (np.tril(A[:M, :M]) if uplo == "L" else np.triu(A[:M, :M])).astype(np.float64)
    X = D[:M, :N].astype(np.float64)
    assert np.isfinite(X).all()
    R = _op(T, trans) @ X - alpha * B[:M, :N].astype(np.float64)
    eps = np.finfo(dt).eps
    berr = np.abs(R).max() / (np.abs(T).sum(axis=1).max() * np.abs(X).max() + abs(alpha) * np.abs(B).max())
    assert berr <= 4 * M * eps, berr


def test_blas3_arguments(lx, dev):
    A = np.zeros((4, 4), order="F")
    f = lx.lib.linompss_dgemm
    assert f(b"X", b"N", 1, 1, 1, 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4) == -1
    assert f(b"N", b"N", 0, 1, 1, 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4) == -3
    assert f(b"N", b"N", 1, 1, 1, 4, 4, 4, 1.0, None, 3, None, 4, 0.0, None, 4) == -11
    assert f(b"N", b"N", 1, 1, 1, 0, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4) == 0
    assert lx.lib.linompss_dsyrk(b"Q", b"N", 1, 4, 4, 1.0, None, 4, 0.0, None, 4) == -1
    assert lx.lib.linompss_dtrsm(b"L", b"L", b"N", b"Z", 4, 4, 1, 1.0, None, 4, None, 4) == -4
    assert lx.lib.linompss_dtrsm(b"R", b"L", b"N", b"N", 4, 5, 1, 1.0, None, 4, None, 4) == -10
    # alpha = 0: B := 0 without reading A (BLAS semantics)
    B = np.ones((4, 3), order="F")
    lx.trsm("L", "L", "N", "N", 4, 3, 0.0, A, B)
    assert (B == 0).all()
