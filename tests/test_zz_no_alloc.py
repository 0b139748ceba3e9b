"""linompss_potrs_dev_noalloc: the solve with an existing factor that
allocates nothing (its workspace is a fixed device array of the library).

Checked through the C-ABI against linompss_potrs_dev (the allocating entry
point) with the same kernels, the 64-row sweeps: bit-identical results, on two
calls in a row of which the second is smaller (the array keeps the first
call's contents), for both element types and for more than 16 right-hand
sides (several launches); sizes past the array's limit are refused before any
device work.  For FP32 from n = 2048 on, linompss_potrs_dev itself takes the
128-row sweeps (a step-matrix pre-pass the fixed array has no room for): the
comparison then forces its 64-row kernels (LINOMPSS_SWEEP2_MIN_N=0), and the
default path is held to FP32 rounding of the same solve.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _factor(lx, orc, n, dtype, seed):
    A = orc.genmat_spd(n, seed, diag_mode=1)
    L = np.asfortranarray(np.linalg.cholesky(A)).astype(dtype, order="F")
    return lx.DeviceMatrix.from_host(L)


def _pair(lx, dL, B, b):
    X1 = lx.DeviceMatrix.from_host(B)
    X2 = lx.DeviceMatrix.from_host(B)
    lx.potrs_dev(dL, X1, b)
    rc = lx.lib.linompss_potrs_dev_noalloc(0 if B.dtype == np.float64 else 1, dL.cols, X2.cols, b, dL.ptr, dL.ld,
                                           X2.ptr, X2.ld, None, None)
    assert rc == 0
    return X1.to_host(), X2.to_host()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_potrs_two_calls(lx, orc, dev, monkeypatch, dtype):
    for n, nr, seed in ((3000, 5, 1), (700, 1, 2)):  # the second call is smaller
        print(f"[noalloc] dtype={np.dtype(dtype).name} n={n} nr={nr}", flush=True)
        dL = _factor(lx, orc, n, dtype, seed)
        B = np.asfortranarray(np.random.default_rng(seed).uniform(-1, 1, (n, nr))).astype(dtype, order="F")
        if dtype == np.float32:  # the default path (128-row sweeps from n = 2048): FP32 rounding apart
            X0, X2 = _pair(lx, dL, B, 256)
            assert np.abs(X0 - X2).max() <= 1e-5 * np.abs(X2).max()
        monkeypatch.setenv("LINOMPSS_SWEEP2_MIN_N", "0")  # linompss_potrs_dev on the 64-row sweeps
        X1, X2 = _pair(lx, dL, B, 256)
        monkeypatch.delenv("LINOMPSS_SWEEP2_MIN_N")
        assert np.array_equal(X1, X2)
        if dtype == np.float64:
            A = orc.genmat_spd(n, seed, diag_mode=1)
            assert orc.backward_error(A, X2, B) <= 1e-14


def test_noalloc_many_rhs_and_full_size(lx, orc, dev):
    """20 right-hand sides (two launches per sweep) at n = 4100 (ragged last
    block), then the largest order the array holds, n = 32768, one column."""
    n, nr = 4100, 20
    print(f"[noalloc] n={n} nr={nr}", flush=True)
    dL = _factor(lx, orc, n, np.float64, 3)
    B = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, (n, nr)))
    X1, X2 = _pair(lx, dL, B, 512)
    assert np.array_equal(X1, X2)
    dL.free()
    n = 32768
    print(f"[noalloc] n={n} nr=1", flush=True)
    # a well conditioned lower factor built on the device: L = chol(A) with A
    # the refinement app's covariance matrix
    dA = lx.gen_covariant(n)
    assert lx.chol_dev(dA, 512) == 0
    B = np.asfortranarray(np.random.default_rng(4).uniform(-1, 1, (n, 1)))
    X1, X2 = _pair(lx, dA, B, 512)
    assert np.array_equal(X1, X2)


def test_noalloc_refuses_past_limit(lx, dev):
    """n > 32768 is refused (LINOMPSS_ERR_LIMIT) and fused_solve is refused
    (-10) before any device memory is touched: the pointers are never read."""
    bogus = ctypes.c_void_p(16)
    assert lx.lib.linompss_potrs_dev_noalloc(0, 32769, 1, 512, bogus, 32769, bogus, 32769, None, None) == -1007
    o = lx.ExecOpts(1, 128, 128, 0, 30.0, 1, 0)
    assert lx.lib.linompss_potrs_dev_noalloc(0, 100, 1, 64, bogus, 100, bogus, 100, None, ctypes.byref(o)) == -10
    assert lx.lib.linompss_potrs_dev_noalloc(2, 100, 1, 64, bogus, 100, bogus, 100, None, None) == -1
