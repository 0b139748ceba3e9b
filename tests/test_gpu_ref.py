"""GPU results against the reference's own path compiled from its sources
(oracle/_ref/libref.so: CHOL_RL / CHOL_LL over its tile tasks, and
dspdsolv_main), on the same inputs, through the C-ABI.

Only the lower triangle is compared (the reference writes into the strict
upper triangle when n % b != 0, chol_rlmain.c:36), with lda = n (it assumes
lda = n, chol_rlmain.c:45,48).  Bars: FP64 factor within 1e-12 relative (Frobenius), FP32
within 1e-5 relative, solutions to a normwise backward error <= 1e-14 and
within 1e-12 of the reference's.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _rel(X, Y):
    return np.linalg.norm(np.tril(X) - np.tril(Y)) / np.linalg.norm(np.tril(Y))


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(256, 64), (700, 128), (1024, 256), (2048, 512)])
def test_dchol_vs_reference(lx, orc, ref, dev, variant, n, b):
    print(f"[gpu vs ref dchol {variant}] n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, 11, diag_mode=1)
    Ag, Ar = A0.copy(order="F"), A0.copy(order="F")
    fn = lx.ompss_dchol_rl if variant == "rl" else lx.ompss_dchol_ll
    assert fn(Ag, b) == 0
    assert getattr(ref, f"ompss_dchol_{variant}")(n, b, b, _p(Ar), n) == 0
    assert _rel(Ag, Ar) <= 1e-12


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(512, 128), (1000, 256)])
def test_schol_vs_reference(lx, orc, ref, dev, variant, n, b):
    print(f"[gpu vs ref schol {variant}] n={n} b={b}", flush=True)
    A0 = orc.genmat_spd(n, 12, diag_mode=1).astype(np.float32, order="F")
    Ag, Ar = A0.copy(order="F"), A0.copy(order="F")
    fn = lx.ompss_schol_rl if variant == "rl" else lx.ompss_schol_ll
    assert fn(Ag, b) == 0
    assert getattr(ref, f"ompss_schol_{variant}")(n, b, b, _p(Ar), n) == 0
    assert _rel(Ag.astype(np.float64), Ar.astype(np.float64)) <= 1e-5


@pytest.mark.parametrize("m,nr,b", [(512, 1, 128), (1024, 3, 256), (4096, 1, 256)])
def test_dspdsolv_vs_reference(lx, orc, ref, dev, m, nr, b):
    print(f"[gpu vs ref dspdsolv] m={m} nr={nr} b={b}", flush=True)
    A0 = orc.genmat_spd(m, 13, diag_mode=1)
    B0 = np.asfortranarray(A0 @ np.ones((m, nr)))
    Ag, Bg = A0.copy(order="F"), B0.copy(order="F")
    Ar, Br = A0.copy(order="F"), B0.copy(order="F")
    assert lx.dspdsolv_main(Ag, Bg, b) == 0
    assert ref.dspdsolv_main(m, nr, b, _p(Ar), _p(Br)) == 0
    assert orc.backward_error(A0, Bg, B0) <= 1e-14
    assert np.abs(Bg - Br).max() <= 1e-12 * np.abs(Br).max()
    assert _rel(Ag, Ar) <= 1e-12
