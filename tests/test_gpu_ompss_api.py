"""The reference's own BLAS-3 and tile-task symbols on the device, through the C-ABI.

``ompss_{d,s}gemm`` (src/gemm/ompss_mm.h:23,27), ``ompss_{d,s}syrk``
(src/syrk/ompss_syrk.h:20,22), ``ompss_{d,s}trsm`` (src/trsm/ompss_trsm.h:20,23)
and the libbblas tile tasks ``task_{d,s}potrf / trsm / syrk / gemm``
(src/blocked_blas/tasks_potrf.h:23, tasks_trsm.h:24, tasks_syrk.h:21,
task_gemm.h:27), called with flags in each of the reference's descriptor
flavours (src/fp/fpmatr.c:14-47): CBLAS values, pointers to one-letter
strings, and the library's own exported OMPSSBLAS_* / OMPSSLAPACK_*
instances.  Results are checked against numpy in FP64 and, where
oracle/_ref is built, against the reference's own task_* compiled from its
sources, called with its own descriptors.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAVOURS = ("enum", "str", "exported")
EXPORTED = {("trans", "N"): "NTRANSP", ("trans", "T"): "TRANSP", ("uplo", "U"): "UPPERTRIANG",
            ("uplo", "L"): "LOWERTRIANG", ("diag", "N"): "NDIAGUNIT", ("diag", "U"): "DIAGUNIT",
            ("side", "L"): "LEFT", ("side", "R"): "RIGHT"}
TOL = {np.float64: 1e-13, np.float32: 2e-5}


def desc(lx, kind, letter, flavour):
    if flavour == "exported":
        return lx.exported_desc("OMPSSBLAS_" + EXPORTED[(kind, letter)])
    return lx.blas_desc(kind, letter, flavour)


def ldesc(lx, letter, flavour):
    if flavour == "exported":
        return lx.exported_desc("OMPSSLAPACK_" + ("UPPERTRIANG" if letter == "U" else "LOWERTRIANG"))
    return lx.lapack_desc(letter, flavour)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pad(rng, rows, cols, ld, dt):
    """A column-major rows x cols matrix inside an ld x cols array whose pad rows hold a marker."""
    M = np.full((ld, cols), 7.25, dtype=dt, order="F")
    M[:rows] = rng.uniform(-1, 1, (rows, cols))
    return M


def _pre(dt):
    return "d" if dt == np.float64 else "s"


def _close(got, ref, bound, dt):
    """|got - ref| <= tol * bound elementwise (bound = |A| |B| style magnitude)."""
    assert np.all(np.abs(got.astype(np.float64) - ref) <= TOL[dt] * (bound + 1e-300))


@pytest.mark.parametrize("flavour", FLAVOURS)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_ompss_gemm_flavours(lx, dev, flavour, dt, ta, tb):
    m, n, k = 130, 97, 75
    print(f"[ompss gemm] {flavour} {_pre(dt)} {ta}{tb}", flush=True)
    rng = np.random.default_rng(hash((ta, tb)) % 1000)
    ar, ac = (k, m) if ta == "T" else (m, k)
    br, bc = (n, k) if tb == "T" else (k, n)
    A, B, C = _pad(rng, ar, ac, ar + 3, dt), _pad(rng, br, bc, br + 1, dt), _pad(rng, m, n, m + 5, dt)
    opA = A[:ar].astype(np.float64).T if ta == "T" else A[:ar].astype(np.float64)
    opB = B[:br].astype(np.float64).T if tb == "T" else B[:br].astype(np.float64)
    alpha, beta = -0.75, 1.5
    ref = alpha * opA @ opB + beta * C[:m].astype(np.float64)
    bound = np.abs(alpha) * np.abs(opA) @ np.abs(opB) + abs(beta) * np.abs(C[:m])
    fn = getattr(lx.lib, f"ompss_{_pre(dt)}gemm")
    rc = fn(desc(lx, "trans", ta, flavour), desc(lx, "trans", tb, flavour), 64, 64, 64, m, n, k, alpha,
            _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(C), C.shape[0])
    assert rc == 0
    _close(C[:m], ref, bound, dt)
    assert np.all(C[m:] == 7.25)  # rows past m in the leading dimension untouched


@pytest.mark.parametrize("flavour", FLAVOURS)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("L", "T"), ("U", "N"), ("U", "T")])
def test_ompss_syrk_flavours(lx, dev, flavour, dt, uplo, trans):
    n, k = 111, 64
    print(f"[ompss syrk] {flavour} {_pre(dt)} {uplo}{trans}", flush=True)
    rng = np.random.default_rng(5)
    ar, ac = (k, n) if trans == "T" else (n, k)
    A, C = _pad(rng, ar, ac, ar + 2, dt), _pad(rng, n, n, n + 1, dt)
    C0 = C.copy(order="F")
    opA = A[:ar].astype(np.float64).T if trans == "T" else A[:ar].astype(np.float64)
    alpha, beta = 0.5, -2.0
    ref = alpha * opA @ opA.T + beta * C0[:n].astype(np.float64)
    bound = abs(alpha) * np.abs(opA) @ np.abs(opA.T) + abs(beta) * np.abs(C0[:n])
    getattr(lx.lib, f"ompss_{_pre(dt)}syrk")(desc(lx, "uplo", uplo, flavour), desc(lx, "trans", trans, flavour),
                                             32, n, k, alpha, _p(A), A.shape[0], beta, _p(C), C.shape[0])
    tri = np.tril_indices(n) if uplo == "L" else np.triu_indices(n)
    other = np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)
    assert np.all(np.abs(C[:n][tri] - ref[tri]) <= TOL[dt] * bound[tri])
    assert np.array_equal(C[:n][other], C0[:n][other])  # the other triangle is not referenced


def _tri(rng, na, uplo, dt):
    """A triangular matrix with a dominant diagonal (condition ~ 1)."""
    T = rng.uniform(-1, 1, (na, na)) / na
    T = np.tril(T) if uplo == "L" else np.triu(T)
    T[np.diag_indices(na)] = rng.uniform(1, 2, na)
    return T


@pytest.mark.parametrize("flavour", FLAVOURS)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("side,uplo,trans,diag", [("L", "L", "N", "N"), ("L", "U", "T", "U"),
                                                  ("R", "L", "T", "N"), ("R", "U", "N", "U"),
                                                  ("L", "L", "T", "N"), ("R", "L", "N", "U")])
def test_ompss_trsm_flavours(lx, dev, flavour, dt, side, uplo, trans, diag):
    m, n = 120, 77
    print(f"[ompss trsm] {flavour} {_pre(dt)} {side}{uplo}{trans}{diag}", flush=True)
    rng = np.random.default_rng(9)
    na = m if side == "L" else n
    T = _tri(rng, na, uplo, np.float64)
    A = np.full((na + 4, na), np.nan, dtype=dt, order="F")  # the other triangle holds NaN: never read
    A[:na][np.tril_indices(na) if uplo == "L" else np.triu_indices(na)] = \
        T[np.tril_indices(na) if uplo == "L" else np.triu_indices(na)]
    Teff = T.astype(dt).astype(np.float64)
    if diag == "U":
        Teff[np.diag_indices(na)] = 1.0
        A[np.arange(na), np.arange(na)] = 1e30  # the diagonal is not read either
    op = Teff.T if trans == "T" else Teff
    B = _pad(rng, m, n, m + 2, dt)
    alpha = 1.25
    ref = np.linalg.solve(op, alpha * B[:m].astype(np.float64)) if side == "L" else \
        np.linalg.solve(op.T, alpha * B[:m].astype(np.float64).T).T
    getattr(lx.lib, f"ompss_{_pre(dt)}trsm")(desc(lx, "side", side, flavour), desc(lx, "uplo", uplo, flavour),
                                             desc(lx, "trans", trans, flavour), desc(lx, "diag", diag, flavour),
                                             m, n, 32, alpha, _p(A), A.shape[0], _p(B), B.shape[0])
    assert np.all(np.isfinite(B[:m]))
    assert np.abs(B[:m] - ref).max() <= 50 * TOL[dt] * np.abs(ref).max()
    assert np.all(B[m:] == 7.25)


@pytest.mark.parametrize("flavour", FLAVOURS)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 64, 200, 700])
def test_task_potrf_flavours(lx, orc, dev, flavour, dt, uplo, n):
    """task_{d,s}potrf factors the uplo triangle in place (LAPACK potrf) and
    leaves the other triangle and the pad rows alone."""
    print(f"[task potrf] {flavour} {_pre(dt)} {uplo} n={n}", flush=True)
    S = orc.genmat_spd(n, 3, diag_mode=1)
    lda = n + 3
    A = np.full((lda, n), 7.25, dtype=dt, order="F")
    A[:n] = S
    A0 = A.copy(order="F")
    getattr(lx.lib, f"task_{_pre(dt)}potrf")(ldesc(lx, uplo, flavour), n, _p(A), lda, 0)
    L = np.linalg.cholesky(S.astype(dt).astype(np.float64))
    got = np.tril(A[:n]) if uplo == "L" else np.triu(A[:n]).T
    assert np.abs(got - L).max() <= 20 * TOL[dt] * np.abs(L).max() * max(1.0, np.sqrt(n) / 8)
    other = np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)
    assert np.array_equal(A[:n][other], A0[:n][other])
    assert np.all(A[n:] == 7.25)


@pytest.mark.parametrize("flavour", FLAVOURS)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_tile_tasks_one_step(lx, orc, dev, flavour, dt):
    """One right-looking step of the reference's CHOL_RL on a 3 x 3 grid of
    b x b tiles, spelled with the tile tasks as chol_rlmain.c:22-53 spells it
    (potrf of the diagonal tile, trsm R/L/T/N of the panel, syrk L/N and gemm
    N/T of the trailing tiles), then the rest of the factor by task_potrf of
    the trailing matrix: the result is the Cholesky factor."""
    b, nt = 96, 3
    n = b * nt
    print(f"[tile tasks] {flavour} {_pre(dt)} n={n} b={b}", flush=True)
    S = orc.genmat_spd(n, 11, diag_mode=1)
    A = np.asfortranarray(S.astype(dt))
    pre = _pre(dt)
    L, R, T, N = (desc(lx, "uplo", "L", flavour), desc(lx, "side", "R", flavour),
                  desc(lx, "trans", "T", flavour), desc(lx, "trans", "N", flavour))
    ND = desc(lx, "diag", "N", flavour)
    lib = lx.lib

    def tile(i, j):
        return ctypes.c_void_p(A.ctypes.data + (j * b * n + i * b) * A.itemsize)

    getattr(lib, f"task_{pre}potrf")(ldesc(lx, "L", flavour), b, tile(0, 0), n, 0)
    for i in range(1, nt):
        getattr(lib, f"task_{pre}trsm")(R, L, T, ND, b, b, 1.0, tile(0, 0), n, tile(i, 0), n, 0)
    for i in range(1, nt):
        getattr(lib, f"task_{pre}syrk")(L, N, b, b, -1.0, tile(i, 0), n, 1.0, tile(i, i), n, 0)
        for j in range(1, i):
            getattr(lib, f"task_{pre}gemm")(N, T, b, b, b, -1.0, tile(i, 0), n, tile(j, 0), n, 1.0, tile(i, j),
                                            n, 0)
    getattr(lib, f"task_{pre}potrf")(ldesc(lx, "L", flavour), n - b, tile(1, 1), n, 0)
    Lref = np.linalg.cholesky(S.astype(dt).astype(np.float64))
    assert np.abs(np.tril(A) - Lref).max() <= 50 * TOL[dt] * np.abs(Lref).max()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_tile_tasks_match_reference_tasks(lx, ref, dev, dt):
    """The same task calls on the reference's own task_* (oracle/_ref: its
    tasks_*.c over MKL, Fortran-BLAS descriptors it exports) and on this
    library's, each with its own descriptor instances: the tiles agree to
    rounding."""
    b = 80
    pre = _pre(dt)
    print(f"[tasks vs reference] {pre} b={b}", flush=True)
    rng = np.random.default_rng(2)
    G = rng.uniform(-1, 1, (b, b))
    S = G @ G.T / b + np.eye(b)
    X = np.asfortranarray(rng.uniform(-1, 1, (b, b)).astype(dt))
    Y = np.asfortranarray(rng.uniform(-1, 1, (b, b)).astype(dt))
    out = []
    for L_ in (lx.lib, ref):
        def d(name, lib=L_):
            return ctypes.c_uint64.in_dll(lib, name).value
        for nm, args in (("potrf", 5), ("trsm", 12), ("syrk", 11), ("gemm", 14)):
            getattr(L_, f"task_{pre}{nm}").restype = None
        fp = ctypes.c_double if dt == np.float64 else ctypes.c_float
        u64, ci, vp = ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        getattr(L_, f"task_{pre}potrf").argtypes = [u64, ci, vp, ci, ci]
        getattr(L_, f"task_{pre}trsm").argtypes = [u64] * 4 + [ci, ci, fp, vp, ci, vp, ci, ci]
        getattr(L_, f"task_{pre}syrk").argtypes = [u64, u64, ci, ci, fp, vp, ci, fp, vp, ci, ci]
        getattr(L_, f"task_{pre}gemm").argtypes = [u64, u64, ci, ci, ci, fp, vp, ci, vp, ci, fp, vp, ci, ci]
        A = np.asfortranarray(S.astype(dt))
        P, Q, C = X.copy(order="F"), Y.copy(order="F"), np.asfortranarray(S.astype(dt))
        getattr(L_, f"task_{pre}potrf")(d("OMPSSLAPACK_LOWERTRIANG"), b, _p(A), b, 0)
        getattr(L_, f"task_{pre}trsm")(d("OMPSSBLAS_RIGHT"), d("OMPSSBLAS_LOWERTRIANG"), d("OMPSSBLAS_TRANSP"),
                                       d("OMPSSBLAS_NDIAGUNIT"), b, b, 1.0, _p(A), b, _p(P), b, 0)
        getattr(L_, f"task_{pre}syrk")(d("OMPSSBLAS_LOWERTRIANG"), d("OMPSSBLAS_NTRANSP"), b, b, -1.0, _p(P), b,
                                       1.0, _p(C), b, 0)
        getattr(L_, f"task_{pre}gemm")(d("OMPSSBLAS_NTRANSP"), d("OMPSSBLAS_TRANSP"), b, b, b, -1.0, _p(P), b,
                                       _p(Q), b, 1.0, _p(C), b, 0)
        out.append((np.tril(A), P, C))
    for g, r in zip(out[0], out[1]):
        assert np.abs(g.astype(np.float64) - r).max() <= 100 * TOL[dt] * max(1.0, np.abs(r).max())
