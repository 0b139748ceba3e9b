"""The library's kernel source executed on the CPU (tests/cpu_emu).

The persistent executor and every tile kernel are compiled for the host with
an emulation of the HIP device environment (one workgroup of 256 threads,
timed barriers, emulated MFMA and shuffles) and run on small plans; results
are checked against the oracle and LAPACK.  This pins the kernels' indexing,
barrier matching and dependency waits without a GPU.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

EMU = Path(__file__).resolve().parent / "cpu_emu"


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", str(EMU)], check=True)
    E = ctypes.CDLL(str(EMU / "_build" / "liblx_emu.so"))
    E.lx_emu_run.restype = ctypes.c_int
    E.lx_emu_run.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
                                                   ctypes.c_longlong, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3

    def run(dtype, what, n, b, nrhs, A, X=None, policy=1):
        info, done, nu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = E.lx_emu_run(dtype, what, n, b, nrhs, A.ctypes.data_as(ctypes.c_void_p), A.shape[0],
                          X.ctypes.data_as(ctypes.c_void_p) if X is not None else None,
                          X.shape[0] if X is not None else 0, policy, ctypes.byref(info), ctypes.byref(done),
                          ctypes.byref(nu))
        assert rc == 0
        assert done.value == nu.value
        return info.value
    return run


@pytest.mark.parametrize("what", [0, 1])
@pytest.mark.parametrize("n,b", [(1, 1), (8, 4), (37, 8), (200, 256)])
def test_factor_fp64(emu, orc, what, n, b):
    A0 = orc.genmat_spd(n, 0, diag_mode=1)
    A = A0.copy(order="F")
    assert emu(0, what, n, b, 0, A) == 0
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(A) - Lref) / np.linalg.norm(Lref) <= 1e-13
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))


def test_factor_fp32_split_bf16_updates(emu, orc):
    """The FP32 factor with its tile-loop SYRK/GEMM units on the split-BF16
    unit (lx_unit_gemm_bx6): FP32-level agreement with LAPACK in FP64."""
    E = ctypes.CDLL(str(EMU / "_build" / "liblx_emu.so"))
    n, b = 300, 128
    A0 = orc.genmat_spd(n, 4, diag_mode=1)
    Af = A0.astype(np.float32, order="F")
    E.lx_emu_set_update_prec(1)
    try:
        assert emu(1, 0, n, b, 0, Af) == 0
    finally:
        E.lx_emu_set_update_prec(0)
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(Af) - Lref) / np.linalg.norm(Lref) <= 1e-6


def test_factor_fp32_and_policy0(emu, orc):
    n, b = 97, 32
    A0 = orc.genmat_spd(n, 2, diag_mode=1)
    Af = A0.astype(np.float32, order="F")
    assert emu(1, 0, n, b, 0, Af, policy=0) == 0
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(Af) - Lref) / np.linalg.norm(Lref) <= 1e-6


@pytest.mark.parametrize("n,b,nr", [(37, 8, 2)])
def test_posv_single_launch(emu, orc, n, b, nr):
    A0 = orc.genmat_spd(n, 0, diag_mode=1)
    B0 = np.asfortranarray(A0 @ np.ones((n, nr)))
    A, X = A0.copy(order="F"), B0.copy(order="F")
    assert emu(0, 2, n, b, nr, A, X) == 0
    assert np.abs(X - 1.0).max() <= 1e-13


def test_sweeps_match_oracle(emu, orc):
    n, b, nr = 70, 16, 3
    rng = np.random.default_rng(3)
    L = np.asfortranarray(np.tril(rng.uniform(0, 1, (n, n))) + n * np.eye(n))
    B0 = np.asfortranarray(rng.uniform(-1, 1, (n, nr)))
    for what, trans in ((4, 0), (5, 1)):
        X = B0.copy(order="F")
        emu(0, what, n, b, nr, L.copy(order="F"), X)
        Xo = B0.copy(order="F")
        orc.ltrsm_lower(trans, L, Xo, b)
        assert np.abs(X - Xo).max() <= 1e-13 * np.abs(Xo).max()


def test_info_on_indefinite(emu, orc):
    n, b = 50, 25  # the bad pivot (global column 31) is local column 6 of tile 1
    A = orc.genmat_spd(n, 1, diag_mode=1)
    A[30, 30] = -5.0
    assert emu(0, 0, n, b, 0, A.copy(order="F")) == 31
    assert emu(0, 1, n, b, 0, A.copy(order="F")) == 31


@pytest.fixture(scope="module")
def emu_sweep(emu):
    E = ctypes.CDLL(str(EMU / "_build" / "liblx_emu.so"))
    E.lx_emu_sweep.restype = ctypes.c_int
    E.lx_emu_sweep.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
                                                   ctypes.c_longlong]

    def run(trans, L, X):
        dtype = 0 if L.dtype == np.float64 else 1
        rc = E.lx_emu_sweep(dtype, trans, L.shape[0], X.shape[1], L.ctypes.data_as(ctypes.c_void_p), L.shape[0],
                            X.ctypes.data_as(ctypes.c_void_p), X.shape[0])
        assert rc == 0
    return run


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,nr", [(1, 1), (100, 3), (200, 16), (700, 2), (300, 7)])
def test_sweep_kernels(emu_sweep, orc, dtype, n, nr):
    """lx_sweep.hip's forward and backward kernels against the oracle's blocked sweeps."""
    rng = np.random.default_rng(n + nr)
    L = np.asfortranarray(np.tril(rng.uniform(0, 1, (n, n))) + n * np.eye(n)).astype(dtype, order="F")
    B0 = np.asfortranarray(rng.uniform(-1, 1, (n, nr))).astype(dtype, order="F")
    for trans in (0, 1):
        X = B0.copy(order="F")
        emu_sweep(trans, L, X)
        Xo = B0.astype(np.float64, order="F")
        orc.ltrsm_lower(trans, np.asfortranarray(L.astype(np.float64)), Xo, 32)
        tol = 1e-13 if dtype == np.float64 else 1e-5
        assert np.abs(X - Xo).max() <= tol * np.abs(Xo).max()
