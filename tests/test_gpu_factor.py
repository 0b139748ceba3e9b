"""GPU parity of the tiled factor (ompss_{d,s}chol_{rl,ll}) against the oracle.

Reference behaviour: src/cholesky/dense/chol_rlmain.c:22-53 and
chol_llmain.c:23-57 (lower factor in place, strict upper triangle untouched),
checked like tests/dense_tests/potrf_test.c:6-37 (||L L^T - A||/(||A|| N eps)
< 10) and against the unique Cholesky factor (relative error <= 1e-12 for
FP64, the plain path's bar).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (n, b): square, ragged last tile, tile larger than n, the reference
# unit test's 1989/533, and several tile sizes of the sweep in tests/dpotrf.cfg.
CASES = [(1, 1), (8, 4), (37, 8), (64, 16), (130, 32), (256, 64), (300, 128), (533, 1024),
         (1024, 128), (1024, 256), (1024, 384), (1989, 533), (2048, 512)]


def _spd_gauss(n, seed, alpha):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + alpha * np.eye(n)
    return np.asfortranarray(A)


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", CASES)
def test_dchol_matches_lapack(lx, orc, dev, n, b, variant):
    print(f"[dchol {variant}] n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, seed=0, diag_mode=1)
    A = A0.copy(order="F")
    fn = lx.ompss_dchol_rl if variant == "rl" else lx.ompss_dchol_ll
    assert fn(A, b) == 0
    Lref = np.linalg.cholesky(A0)
    rel = np.linalg.norm(np.tril(A) - Lref) / np.linalg.norm(Lref)
    assert rel <= 1e-12, rel
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1)), "strict upper triangle modified"
    assert orc.potrf_resid(A0, A) < 10.0


@pytest.mark.parametrize("n,b", [(96, 32), (500, 128), (1000, 256)])
def test_dchol_vs_oracle_illconditioned(lx, orc, dev, n, b):
    print(f"[dchol gauss] n={n} b={b}", flush=True)
    A0 = _spd_gauss(n, 7, 1e-3)
    A = A0.copy(order="F")
    assert lx.ompss_dchol_rl(A, b) == 0
    Lo = A0.copy(order="F")
    assert orc.chol(Lo, b, "rl") == 0
    rel = np.linalg.norm(np.tril(A) - np.tril(Lo)) / np.linalg.norm(np.tril(Lo))
    # cond(A) ~ 4e3: the factors agree to a few ulps times sqrt(cond)
    assert rel <= 1e-12, rel
    assert orc.potrf_resid(A0, A) < 10.0


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(8, 4), (130, 32), (533, 128), (1989, 533), (2048, 512)])
def test_schol(lx, orc, dev, n, b, variant):
    print(f"[schol {variant}] n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, seed=3, diag_mode=1).astype(np.float32, order="F")
    A = A0.copy(order="F")
    fn = lx.ompss_schol_rl if variant == "rl" else lx.ompss_schol_ll
    assert fn(A, b) == 0
    Lo = A0.copy(order="F")
    assert orc.chol(Lo, b, variant) == 0
    Lref = np.linalg.cholesky(A0.astype(np.float64))
    rel = np.linalg.norm(np.tril(A).astype(np.float64) - Lref) / np.linalg.norm(Lref)
    relo = np.linalg.norm(np.tril(Lo).astype(np.float64) - Lref) / np.linalg.norm(Lref)
    assert rel <= 4 * max(relo, 1e-7), (rel, relo)
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))


@pytest.mark.parametrize("n,b,k", [(64, 16, 20), (300, 64, 150), (1000, 256, 700)])
def test_not_spd_reports_info(lx, dev, n, b, k):
    """A matrix whose leading minor of order k+1 is indefinite: LAPACK info = k+1."""
    print(f"[info] n={n} b={b} k={k}", flush=True)
    A = _spd_gauss(n, 11, 1.0)
    A[k, k] = -abs(A[k, k]) - 10.0 * n
    A = np.asfortranarray(A)
    assert lx.ompss_dchol_rl(A.copy(order="F"), b) == k + 1
    assert lx.ompss_dchol_ll(A.copy(order="F"), b) == k + 1


def test_lda_larger_than_n(lx, orc, dev):
    n, lda, b = 200, 257, 64
    print("[lda]", flush=True)
    A0 = orc.genmat_spd(n, seed=5, diag_mode=0, lda=lda)
    A = A0.copy(order="F")
    assert lx.ompss_dchol_rl(A, b, n=n, lda=lda) == 0
    Lref = np.linalg.cholesky(A0[:n, :n])
    assert np.linalg.norm(np.tril(A[:n]) - Lref) / np.linalg.norm(Lref) <= 1e-12
    assert np.array_equal(A[n:], A0[n:]), "rows beyond n modified"


def test_device_api_factor(lx, orc, dev):
    n, b = 1536, 256
    print("[dev api]", flush=True)
    A0 = orc.genmat_spd(n, seed=9, diag_mode=1)
    dA = lx.DeviceMatrix.from_host(A0)
    assert lx.chol_dev(dA, b) == 0
    L = np.tril(dA.to_host())
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) <= 1e-12
    st = lx.last_exec_stats()
    assert st.units > 0 and st.exec_ms > 0


@pytest.mark.parametrize("n,b", [(700, 128), (2048, 512), (1989, 533)])
def test_schol_split_bf16_updates(lx, orc, dev, n, b):
    """FP32 factor with the tile loop's SYRK/GEMM on split-BF16 matrix cores
    (update_prec = 1): FP32-level agreement with LAPACK (FP64), like the FP32
    matrix-core path."""
    print(f"[schol bf16x6] n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, seed=6, diag_mode=1)
    dA = lx.DeviceMatrix.from_host(A0.astype(np.float32, order="F"))
    opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, 1)
    assert lx.chol_dev(dA, b, "rl", opts) == 0
    L = np.tril(dA.to_host().astype(np.float64))
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) <= 1e-5
    dA.free()


@pytest.mark.parametrize("n,b", [(700, 128), (2048, 512), (1989, 533)])
def test_schol_split_fp16_updates(lx, orc, dev, n, b):
    """FP32 factor with the tile loop's SYRK/GEMM on split-FP16 matrix cores
    (update_prec = 2) on a matrix whose diagonal is below 1, as the
    refinement driver equilibrates it: FP32-level agreement with LAPACK."""
    print(f"[schol fp16x3] n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, seed=6, diag_mode=1) / 2.0 ** np.ceil(np.log2(n))
    dA = lx.DeviceMatrix.from_host(A0.astype(np.float32, order="F"))
    opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, 2)
    assert lx.chol_dev(dA, b, "rl", opts) == 0
    L = np.tril(dA.to_host().astype(np.float64))
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) <= 1e-5
    dA.free()


def test_schol_split_fp16_out_of_range_reports(lx, orc, dev):
    """Panel entries past FP16's range (x 2^12) end in a non-positive pivot
    (info > 0), never in a silently wrong factor."""
    n, b = 700, 128
    A0 = orc.genmat_spd(n, seed=6, diag_mode=1) / 1024.0
    S = np.where(np.arange(n) >= b, 2.0 ** 16, 1.0)
    dA = lx.DeviceMatrix.from_host((S[:, None] * A0 * S[None, :]).astype(np.float32, order="F"))
    opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, 2)
    assert lx.chol_dev(dA, b, "rl", opts) > 0
    dA.free()


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(700, 128), (4100, 512), (1989, 533), (3000, 256)])
def test_schol_fp16_planes_bit_identical(lx, orc, dev, monkeypatch, variant, n, b):
    """The split-FP16 factor with its GEMM / SYRK operands read from the FP16
    planes the TRSM units write equals the factor that splits the FP32
    operands in place (LINOMPSS_PLANES=0) bit for bit: the same values are
    split the same way and multiplied in the same order.  b = 533 is not a
    multiple of 16, so no unit can use the planes there; n = 4100 ends on a
    ragged panel of 4 columns."""
    print(f"[schol planes] n={n} b={b} {variant}", flush=True)
    A0 = orc.genmat_spd(n, seed=7, diag_mode=1) / 2.0 ** np.ceil(np.log2(n))
    opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, 2)
    out = []
    for planes in ("1", "0"):
        monkeypatch.setenv("LINOMPSS_PLANES", planes)
        dA = lx.DeviceMatrix.from_host(A0.astype(np.float32, order="F"))
        assert lx.chol_dev(dA, b, variant, opts) == 0
        out.append(np.tril(dA.to_host()))
        dA.free()
    assert np.array_equal(out[0], out[1])
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(out[0] - Lref) / np.linalg.norm(Lref) <= 1e-5


@pytest.mark.parametrize("dtype,variant,n,b", [(np.float64, "rl", 4096, 512), (np.float64, "rl", 3000, 256),
                                               (np.float32, "rl", 4096, 512), (np.float64, "ll", 4096, 512)])
def test_chol_tall_units(lx, orc, dev, monkeypatch, dtype, variant, n, b):
    """GEMM units of up to a whole tile's rows (128-row blocks under one
    publication) and 128-row panel solves, forced at every step
    (LINOMPSS_WIDE_UNITS=0; by default only steps with two units per
    workgroup left use them, which small orders never reach): the factor
    equals the factor of 128-row units bit for bit (the same blocks with the
    same arithmetic) and agrees with LAPACK; FP32 with the split-FP16 updates."""
    print(f"[tall units] {np.dtype(dtype).name} {variant} n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, seed=9, diag_mode=1) / 2.0 ** np.ceil(np.log2(n))
    opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, 2 if dtype == np.float32 else 0)
    out = []
    for env in ({"LINOMPSS_WIDE_UNITS": "0", "LINOMPSS_WIDE_M": "512", "LINOMPSS_TRSM_ROWS": "128"},
                {"LINOMPSS_WIDE_M": "128", "LINOMPSS_TRSM_ROWS": "64"}):
        for k in ("LINOMPSS_WIDE_UNITS", "LINOMPSS_WIDE_M", "LINOMPSS_TRSM_ROWS"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        dA = lx.DeviceMatrix.from_host(A0.astype(dtype, order="F"))
        assert lx.chol_dev(dA, b, variant, opts) == 0
        out.append(np.tril(dA.to_host()))
        dA.free()
    assert np.array_equal(out[0], out[1])
    Lref = np.linalg.cholesky(A0)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    assert np.linalg.norm(out[0] - Lref) / np.linalg.norm(Lref) <= tol


@pytest.mark.parametrize("dtype,n,b,g", [(np.float64, 4096, 256, 2), (np.float64, 3000, 256, 3),
                                         (np.float64, 2100, 128, 4), (np.float32, 4096, 256, 2),
                                         (np.float32, 3001, 512, 3), (np.float32, 4100, 256, 4)])
def test_chol_grouped_updates(lx, orc, dev, monkeypatch, dtype, n, b, g):
    """Grouped trailing updates (tiles right of a group of g steps take the
    group's g panels in one K = g b pass; g = 2 is the default) against one
    pass per panel (LINOMPSS_UPDATE_GROUP=1) and LAPACK, on exact and ragged
    tilings; FP32 with the split-FP16 updates from the planes."""
    print(f"[grouped] {np.dtype(dtype).name} n={n} b={b} g={g}", flush=True)
    A0 = orc.genmat_spd(n, seed=10, diag_mode=1) / 2.0 ** np.ceil(np.log2(n))
    opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, 2 if dtype == np.float32 else 0)
    out, units = [], []
    for grp in (str(g), "1"):
        monkeypatch.setenv("LINOMPSS_UPDATE_GROUP", grp)
        dA = lx.DeviceMatrix.from_host(A0.astype(dtype, order="F"))
        assert lx.chol_dev(dA, b, "rl", opts) == 0
        units.append(lx.last_exec_stats().units)
        out.append(dA.to_host())
        dA.free()
    assert units[0] < units[1]
    Lref = np.linalg.cholesky(A0)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for o in out:
        assert np.linalg.norm(np.tril(o) - Lref) / np.linalg.norm(Lref) <= tol
        assert np.array_equal(np.triu(o, 1), np.triu(A0.astype(dtype), 1))
