"""The oracle against the reference's own path, compiled from its sources.

``oracle/_ref/libref.so`` is the reference's CHOL_RL / CHOL_LL loop nests, its
tile tasks (tasks_potrf.c, tasks_trsm.c, tasks_syrk.c, task_gemm.c), its entry
wrappers (ompss_dense_chol.c) and dspdsolv_main (spdsolv_main.c, with the
header-inline bblas_dltrsm sweeps), built by ``make -C oracle ref`` straight
from the reference tree: sequentially (the OmpSs task pragmas are ignored, so
program order is a valid order of the task DAG) over the same sequential MKL
the oracle calls.  Every tile then receives the same BLAS calls in the same
order in both, so the oracle must reproduce the reference bit for bit.

Inputs use lda = n and tile-multiple or ragged n: the reference assumes
lda = n (its GEMM/SYRK targets use n as the leading dimension,
chol_rlmain.c:45,48) and, for n % b != 0, its RL TRSM row count runs past the
bottom into the strict upper triangle (chol_rlmain.c:36), so only the lower
triangle is compared.
"""
import ctypes

import numpy as np
import pytest


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(64, 16), (256, 64), (300, 64), (512, 128), (700, 256)])
def test_dchol_bit_exact(ref, orc, variant, n, b):
    A0 = orc.genmat_spd(n, 7, diag_mode=1)
    Ar, Ao = A0.copy(order="F"), A0.copy(order="F")
    assert getattr(ref, f"ompss_dchol_{variant}")(n, b, b, _p(Ar), n) == 0
    assert orc.chol(Ao, b, variant) == 0
    assert np.array_equal(np.tril(Ar), np.tril(Ao))


@pytest.mark.parametrize("variant", ["rl", "ll"])
@pytest.mark.parametrize("n,b", [(128, 32), (400, 96)])
def test_schol_bit_exact(ref, orc, variant, n, b):
    A0 = orc.genmat_spd(n, 8, diag_mode=1).astype(np.float32, order="F")
    Ar, Ao = A0.copy(order="F"), A0.copy(order="F")
    assert getattr(ref, f"ompss_schol_{variant}")(n, b, b, _p(Ar), n) == 0
    assert orc.chol(Ao, b, variant) == 0
    assert np.array_equal(np.tril(Ar), np.tril(Ao))


@pytest.mark.parametrize("m,nr,b", [(256, 1, 64), (512, 3, 128), (768, 2, 256)])
def test_dspdsolv_bit_exact(ref, orc, m, nr, b):
    # spdsolv pads m to a multiple of b in the reference (spdsolv_config.c:24-27)
    A0 = orc.genmat_spd(m, 9, diag_mode=1)
    B0 = np.asfortranarray(A0 @ np.ones((m, nr)))
    Ar, Br = A0.copy(order="F"), B0.copy(order="F")
    Ao, Bo = A0.copy(order="F"), B0.copy(order="F")
    assert ref.dspdsolv_main(m, nr, b, _p(Ar), _p(Br)) == 0
    assert orc.dspdsolv(Ao, Bo, b) == 0
    assert np.array_equal(np.tril(Ar), np.tril(Ao))
    assert np.array_equal(Br, Bo)
    assert orc.backward_error(A0, Br, B0) <= 1e-14
