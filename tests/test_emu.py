"""The library's kernel source executed on the CPU (tests/cpu_emu).

The persistent executor and every tile kernel are compiled for the host with
an emulation of the HIP device environment (one workgroup of 256 threads,
timed barriers, emulated MFMA and shuffles) and run on small plans; results
are checked against the oracle and LAPACK.  This pins the kernels' indexing,
barrier matching and dependency waits without a GPU.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

EMU = Path(__file__).resolve().parent / "cpu_emu"
# LX_EMU_SANITIZE=1 (set by test_emu_sanitized.py, with the sanitizer runtime
# preloaded): the same sources built with AddressSanitizer and
# UndefinedBehaviorSanitizer
SANITIZE = os.environ.get("LX_EMU_SANITIZE") == "1"


def _load_emu():
    subprocess.run(["make", "-s", "-C", str(EMU), "asan" if SANITIZE else "all"], check=True)
    return ctypes.CDLL(str(EMU / "_build" / ("liblx_emu_asan.so" if SANITIZE else "liblx_emu.so")))


@pytest.fixture(scope="module")
def emu():
    E = _load_emu()
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
@pytest.mark.parametrize("n,b", [(1, 1), (8, 4), (37, 16), (200, 256)])
def test_factor_fp64(emu, orc, what, n, b):
    A0 = orc.genmat_spd(n, 0, diag_mode=1)
    A = A0.copy(order="F")
    assert emu(0, what, n, b, 0, A) == 0
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(A) - Lref) / np.linalg.norm(Lref) <= 1e-13
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))


@pytest.mark.parametrize("dtype,n,b,g", [(np.float64, 112, 32, 2), (np.float32, 160, 32, 2),
                                         (np.float64, 150, 32, 3), (np.float64, 160, 32, 4)])
def test_factor_grouped_updates(emu, orc, dtype, n, b, g):
    """RL plans with grouped trailing updates (tiles right of a group of g
    steps take the group's g panels in one K = g b pass): the factor matches
    LAPACK and the plan with one pass per panel, on ragged last tiles too;
    FP32 with its split-FP16 updates read from the planes; the single-launch
    solve over a grouped plan."""
    E = _load_emu()
    A0 = orc.genmat_spd(n, 8, diag_mode=1) / (2.0 * n)
    Lref = np.linalg.cholesky(A0)
    up = 2 if dtype == np.float32 else 0
    dt = 1 if dtype == np.float32 else 0
    E.lx_emu_set_update_prec(up)
    E.lx_emu_set_planes(1 if up else 0)
    try:
        A = A0.astype(dtype, order="F")
        E.lx_emu_set_update_group(g)
        assert emu(dt, 0, n, b, 0, A) == 0
        E.lx_emu_set_update_group(1)
        A1 = A0.astype(dtype, order="F")
        assert emu(dt, 0, n, b, 0, A1) == 0
    finally:
        E.lx_emu_set_update_group(1)
        E.lx_emu_set_planes(0)
        E.lx_emu_set_update_prec(0)
    tol = 1e-13 if dtype == np.float64 else 1e-6
    assert np.linalg.norm(np.tril(A) - Lref) / np.linalg.norm(Lref) <= tol
    assert np.linalg.norm(np.tril(A) - np.tril(A1)) / np.linalg.norm(Lref) <= tol
    assert np.array_equal(np.triu(A, 1), np.triu(A0.astype(dtype), 1))
    if dtype == np.float64 and g == 2:
        # the fused factor + sweeps launch over a grouped plan
        X = np.asfortranarray(A0 @ np.ones((n, 1)))
        A2 = A0.copy(order="F")
        E.lx_emu_set_update_group(g)
        try:
            assert emu(0, 2, n, b, 1, A2, X) == 0
        finally:
            E.lx_emu_set_update_group(1)
        assert np.abs(X - 1.0).max() <= 1e-10


def test_factor_fp32_split_bf16_updates(emu, orc):
    """The FP32 factor with its tile-loop SYRK/GEMM units on the split-BF16
    unit bodies: FP32-level agreement with LAPACK in FP64."""
    E = _load_emu()
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


@pytest.mark.parametrize("flags,M,N,K,gdiag", [
    (6, 128, 256, 64, 0),    # one full 128 x 256 split-BF16 block
    (6, 100, 512, 40, 0),    # two wide blocks, ragged M and K
    (6, 128, 384, 48, 0),    # a wide block, then a 128 x 128 block
    (6, 77, 200, 33, 0),     # one ragged wide block
    (6, 128, 300, 32, 0),    # a wide block, then a 44-column remainder
    (7, 128, 256, 32, 100),  # lower mask across the wide block
    (3, 128, 128, 32, 0),    # SYRK tile on the 128 x 128 body
    (2, 100, 300, 40, 0),    # 128 x 128 blocks only
    (14, 128, 256, 64, 0),   # the FP16 split: one full wide block
    (14, 77, 200, 33, 0),    # ragged wide block
    (14, 128, 384, 48, 0),   # wide block, then a 128 x 128 block
    (15, 128, 256, 32, 100), # lower mask
    (11, 128, 128, 32, 0),   # SYRK tile on the 128 x 128 body
    (10, 100, 300, 40, 0),   # 128 x 128 blocks only
])
def test_gemm_unit_split_bf16(emu, flags, M, N, K, gdiag):
    """One GEMM / SYRK unit through the executor's dispatch (lx_run_unit) on the
    split bodies (BF16, or FP16 with bit 3): C -= A B^T within FP32
    accumulation error (plus the split's error per product) of the FP64
    product, entries outside the lower mask and rows / columns past M / N
    untouched."""
    E = _load_emu()
    F, I, L, P = ctypes.c_float, ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p
    E.lx_emu_gemm_unit.argtypes = [I, I, I, I, P, L, P, L, I]
    rng = np.random.default_rng(M * 1000 + N + K)
    Pm = np.asfortranarray(rng.standard_normal((M + N, K)).astype(np.float32))
    ldc = M + 5
    C0 = np.asfortranarray(rng.standard_normal((ldc, N)).astype(np.float32))
    C = C0.copy(order="F")
    E.lx_emu_gemm_unit(flags, M, N, K, C.ctypes.data_as(P), ldc, Pm.ctypes.data_as(P), M + N, gdiag)
    A64, B64 = Pm[:M].astype(np.float64), Pm[M:].astype(np.float64)
    ref = C0[:M].astype(np.float64) - A64 @ B64.T
    mask = np.ones((M, N), dtype=bool)
    if flags & 1:
        mask = np.arange(M)[:, None] + gdiag >= np.arange(N)[None, :]
    # FP32 accumulation, plus the split's own error per product: below 2^-24
    # for the BF16 split, 3 * 2^-22 for the FP16 one (bit 3)
    per = 3 * 2.0 ** -22 if flags & 8 else 0.0
    tol = (4 * K * 2.0 ** -24 + per) * (np.abs(A64) @ np.abs(B64).T).max() + 2.0 ** -23 * np.abs(ref).max()
    assert np.abs(C[:M][mask] - ref[mask]).max() <= tol
    assert np.array_equal(C[:M][~mask], C0[:M][~mask])
    assert np.array_equal(C[M:], C0[M:])


def test_factor_fp32_split_fp16_updates(emu, orc):
    """The FP32 factor on the split-FP16 bodies: on a matrix with its diagonal
    below 1 (as the refinement driver equilibrates it) FP32-level agreement
    with LAPACK; with entries of the panels the trailing units read past
    FP16's range (x 2^12) the factor reports a non-positive pivot instead of
    a wrong factor."""
    E = _load_emu()
    n, b = 300, 128
    A0 = orc.genmat_spd(n, 4, diag_mode=1) / 512.0  # diagonal 300/512
    Af = A0.astype(np.float32, order="F")
    E.lx_emu_set_update_prec(2)
    try:
        assert emu(1, 0, n, b, 0, Af) == 0
        Lref = np.linalg.cholesky(A0)
        assert np.linalg.norm(np.tril(Af) - Lref) / np.linalg.norm(Lref) <= 2e-6
        # rows past the first tile scaled by 2^16 (S A S, still SPD): their
        # panel entries, the trailing units' operands, reach ~100
        S = np.where(np.arange(n) >= b, 2.0 ** 16, 1.0)
        Big = (S[:, None] * A0 * S[None, :]).astype(np.float32, order="F")
        assert emu(1, 0, n, b, 0, Big) > 0
    finally:
        E.lx_emu_set_update_prec(0)


@pytest.mark.parametrize("flags,M,N,K,gdiag", [
    (10, 128, 512, 64, 0),   # GEMM unit, two 128 x 256 blocks
    (10, 77, 300, 48, 0),    # ragged M and N
    (11, 128, 128, 32, 0),   # SYRK tile
    (11, 96, 200, 16, 40),   # SYRK with the diagonal inside the block
])
def test_gemm_unit_from_planes(emu, flags, M, N, K, gdiag):
    """A split-FP16 unit (bits 1 and 3) with the plane buffer set reads its
    operands from the FP16 planes (the FP32 operands it could also read are
    zeros here): C -= A B^T to the split's accuracy, entries outside the lower
    mask and past M / N untouched."""
    E = _load_emu()
    I, L, P = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p
    E.lx_emu_gemm_unit_planes.argtypes = [I, I, I, I, P, L, P, L, I]
    rng = np.random.default_rng(M * 1000 + N + K)
    Pm = np.asfortranarray(rng.uniform(-1, 1, (M + N, K)).astype(np.float32))
    ldc = M + 5
    C0 = np.asfortranarray(rng.standard_normal((ldc, N)).astype(np.float32))
    C = C0.copy(order="F")
    E.lx_emu_gemm_unit_planes(flags, M, N, K, C.ctypes.data_as(P), ldc, Pm.ctypes.data_as(P), M + N, gdiag)
    A64, B64 = Pm[:M].astype(np.float64), Pm[M:].astype(np.float64)
    ref = C0[:M].astype(np.float64) - A64 @ B64.T
    mask = np.ones((M, N), dtype=bool)
    if flags & 1:
        mask = np.arange(M)[:, None] + gdiag >= np.arange(N)[None, :]
    tol = (4 * K * 2.0 ** -24 + 3 * 2.0 ** -22) * (np.abs(A64) @ np.abs(B64).T).max() + 2.0 ** -23 * np.abs(ref).max()
    assert np.abs(C[:M][mask] - ref[mask]).max() <= tol
    assert np.abs(C[:M][mask] - C0[:M][mask]).max() > 100 * tol  # the planes were read
    assert np.array_equal(C[:M][~mask], C0[:M][~mask])
    assert np.array_equal(C[M:], C0[M:])


@pytest.mark.parametrize("variant,n,b", [(0, 300, 128), (1, 200, 64)])
def test_factor_fp32_fp16_planes(emu, orc, variant, n, b):
    """Split-FP16 trailing updates from the pre-split planes (TRSM units write
    hi and lo of 2^12 L once; GEMM / SYRK units whose operands are whole
    16-column slabs read them, the others split in place): the same FP32
    values split the same way and multiplied in the same order, so the factor
    equals the in-place split's bit for bit, on ragged orders too (n = 200:
    the last panel's 8 columns are not a whole slab)."""
    E = _load_emu()
    A0 = orc.genmat_spd(n, 5, diag_mode=1) / (2.0 * n)
    E.lx_emu_set_update_prec(2)
    try:
        Ap = A0.astype(np.float32, order="F")
        E.lx_emu_set_planes(1)
        assert emu(1, variant, n, b, 0, Ap) == 0
        E.lx_emu_set_planes(0)
        As = A0.astype(np.float32, order="F")
        assert emu(1, variant, n, b, 0, As) == 0
    finally:
        E.lx_emu_set_planes(0)
        E.lx_emu_set_update_prec(0)
    assert np.array_equal(np.tril(Ap), np.tril(As))
    assert np.array_equal(np.triu(Ap, 1), np.triu(A0.astype(np.float32), 1))
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(Ap) - Lref) / np.linalg.norm(Lref) <= 2e-6


def test_factor_fp32_and_policy0(emu, orc):
    n, b = 97, 32
    A0 = orc.genmat_spd(n, 2, diag_mode=1)
    Af = A0.astype(np.float32, order="F")
    assert emu(1, 0, n, b, 0, Af, policy=0) == 0
    Lref = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(Af) - Lref) / np.linalg.norm(Lref) <= 1e-6


@pytest.mark.parametrize("n,b,nr", [(37, 16, 2)])
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
    E = _load_emu()
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
@pytest.mark.parametrize("n,nr", [(1, 1), (100, 3), (200, 16), (700, 2), (300, 7), (520, 16)])
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


@pytest.fixture(scope="module")
def emu_sweep2(emu):
    E = _load_emu()
    E.lx_emu_sweep2.restype = ctypes.c_int
    E.lx_emu_sweep2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong,
                                ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int]

    def run(trans, L, X, C):
        rc = E.lx_emu_sweep2(trans, L.shape[0], X.shape[1], L.ctypes.data_as(ctypes.c_void_p), L.shape[0],
                             X.ctypes.data_as(ctypes.c_void_p), X.shape[0], C)
        assert rc == 0
    return run


@pytest.mark.parametrize("n,nr,C", [(1, 1, 4), (128, 1, 4), (129, 1, 4), (300, 1, 1), (520, 1, 2), (129, 3, 4),
                                    (300, 16, 1)])
def test_sweep2_kernels(emu_sweep2, orc, n, nr, C):
    """The 128-row FP32 sweeps (step matrices from the pre-pass, then the
    one-column kernel, or the 16-column kernel on the emulated matrix cores,
    with C blocks per partial sum) against the oracle's blocked sweeps in
    FP64 on the same FP32 inputs: a Cholesky factor of a Gram matrix plus
    1e-2 I, ragged orders included.  The bound is FP32 rounding of a solve
    with cond(L) of a few hundred (the 64-row kernel measures the same 1e-7 to
    4e-7 on these inputs)."""
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    L = np.asfortranarray(np.linalg.cholesky(G @ G.T / n + 1e-2 * np.eye(n))).astype(np.float32)
    b = np.asfortranarray(rng.uniform(-1, 1, (n + 3, nr)).astype(np.float32))  # ldx = n + 3
    for trans in (0, 1):
        X = b.copy(order="F")
        emu_sweep2(trans, L, X, C)
        Xo = b[:n].astype(np.float64, order="F")
        orc.ltrsm_lower(trans, np.asfortranarray(L.astype(np.float64)), Xo, 32)
        assert np.abs(X[:n] - Xo).max() <= 2e-6 * np.abs(Xo).max()
        assert np.array_equal(X[n:], b[n:])  # rows past n untouched


@pytest.mark.parametrize("nr", [1, 3])
@pytest.mark.parametrize("trans", [0, 1])
def test_sweep2_nan_with_mailbox_bits(emu_sweep2, trans, nr):
    """The 128-row sweeps with an input NaN whose bits equal the 'empty' mark
    of the solved-value and partial-sum mailboxes: it reaches both as the
    canonical quiet NaN, so the sweep ends (rc 0, no abort) with NaN exactly
    where the solution depends on it."""
    n, row = 520, 200
    rng = np.random.default_rng(12)
    L = np.asfortranarray(np.tril(rng.uniform(0.5, 1, (n, n))) + n * np.eye(n)).astype(np.float32)
    X = np.asfortranarray(rng.uniform(-1, 1, (n, nr))).astype(np.float32)
    X.view(np.uint32)[row, 0] = 0xFFFFFFFF
    assert np.isnan(X[row, 0])
    emu_sweep2(trans, L, X, 2)  # asserts rc == 0
    # the 128-row block of `row` is multiplied by its full inverse: all of it turns NaN
    blk = row // 128 * 128
    dep = np.arange(n) >= blk if trans == 0 else np.arange(n) < blk + 128
    assert np.isnan(X[dep, 0]).all()
    assert np.isfinite(X[~dep, 0]).all()
    if nr > 1:
        assert np.isfinite(X[:, 1:]).all()


def _emu_batch(sizes, mats, rhs, max_sweeps=30, tol=1e-14, stall=0.5):
    """lx_batch_ir_kernel on the CPU: one workgroup takes every system."""
    E = _load_emu()
    cnt = len(sizes)
    P = ctypes.c_void_p
    X = [np.zeros(max(int(n), 1)) for n in sizes]
    st = np.zeros(cnt, dtype=[("sweeps", "i4"), ("fallback", "i4"), ("info", "i4"), ("pad", "i4"), ("berr", "f8")])
    order = np.argsort(-np.asarray(sizes), kind="stable").astype(np.int32)
    rc = E.lx_emu_batch(cnt, (ctypes.c_int * cnt)(*[int(n) for n in sizes]),
                        (P * cnt)(*[m.ctypes.data for m in mats]),
                        (ctypes.c_longlong * cnt)(*[max(int(n), 1) for n in sizes]),
                        (P * cnt)(*[b.ctypes.data for b in rhs]), (P * cnt)(*[x.ctypes.data for x in X]),
                        order.ctypes.data_as(P), st.ctypes.data_as(P), max_sweeps, ctypes.c_double(tol),
                        ctypes.c_double(stall))
    assert rc == 0
    return X, st


def test_batch_kernel(orc):
    """The batched refinement kernel (lx_batch.hip) on uneven orders, including
    the empty system, one row, ragged 64-row blocks and the 128-row blocks of
    the in-workgroup factor's trailing update."""
    rng = np.random.default_rng(21)
    sizes = [1, 37, 65, 130, 0]
    mats, rhs = [], []
    for n in sizes:
        G = rng.standard_normal((n, n))
        mats.append(np.asfortranarray(G @ G.T / max(n, 1) + 0.1 * np.eye(n)) if n else np.zeros(1))
        rhs.append(rng.uniform(-1, 1, n) if n else np.zeros(1))
    X, st = _emu_batch(sizes, mats, rhs)
    for i, n in enumerate(sizes):
        if n == 0:
            assert tuple(st[i])[:3] == (0, 0, 0)
            continue
        A, b, x = mats[i], rhs[i], X[i]
        be = orc.backward_error(A, x, b)
        assert st["fallback"][i] == 0 and st["info"][i] == 0
        assert be <= 1e-14 and st["berr"][i] <= 1e-14
        xr = np.linalg.solve(A, b)
        # x - xr = A^-1 (r_xr - r_x): in the infinity norm the two backward
        # errors bound the difference through cond_inf(A)
        xn, rn = np.abs(x).max(), np.abs(xr).max()
        bound = np.linalg.cond(A, np.inf) * (be * xn / rn + orc.backward_error(A, xr, b))
        assert np.abs(x - xr).max() / rn <= 1.0001 * bound


def test_batch_kernel_128_panels(orc):
    """The batched factor in 128-column panels (lx_b_potrf: tall 128-column
    panels factored by the POTRF unit restricted to their columns, one
    trailing update per panel with K = 128) for every size class: a lone
    partial panel (order 90) and a ragged second panel (150); a not-SPD
    system reports its pivot past the first panel.  (Per-class masks and the
    largest orders run on the GPU in tests/test_gpu_batch.py.)"""
    E = _load_emu()
    E.lx_emu_set_batch_panels.argtypes = [ctypes.c_uint]
    mask = 0xFF
    rng = np.random.default_rng(23)
    sizes = [90, 150]
    mats, rhs = [], []
    for n in sizes:
        G = rng.standard_normal((n, n))
        mats.append(np.asfortranarray(G @ G.T / n + 0.1 * np.eye(n)))
        rhs.append(rng.uniform(-1, 1, n))
    bad = mats[1].copy(order="F")
    bad[140, 140] = -1.0  # order 150: column 141 (1-based) is in the second panel
    E.lx_emu_set_batch_panels(mask)
    try:
        X, st = _emu_batch(sizes + [150], mats + [bad], rhs + [rhs[1]])
    finally:
        E.lx_emu_set_batch_panels(0)
    for i, n in enumerate(sizes):
        A, b, x = mats[i], rhs[i], X[i]
        assert st["fallback"][i] == 0 and st["info"][i] == 0
        assert orc.backward_error(A, x, b) <= 1e-14 and st["berr"][i] <= 1e-14
    assert st["fallback"][2] == 1 and st["info"][2] == 141


def test_batch_kernel_fallback(orc):
    """Not positive definite: info from the FP64 factor; too ill-conditioned
    for FP32 refinement: solved by the FP64 factor."""
    rng = np.random.default_rng(22)
    n = 50
    G = rng.standard_normal((n, n))
    A1 = np.asfortranarray(G @ G.T / n + 0.1 * np.eye(n))
    A1[3, 3] = -1.0
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A2 = np.asfortranarray((Q * np.logspace(0, -9, n)) @ Q.T)
    b1, b2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    X, st = _emu_batch([n, n], [A1, A2], [b1, b2])
    assert st["fallback"][0] == 1 and st["info"][0] == 4
    assert st["fallback"][1] == 1 and st["info"][1] == 0
    assert orc.backward_error(A2, X[1], b2) <= 1e-14


def _emu_lib():
    E = _load_emu()
    P, I, L, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    E.lx_emu_adj_diag.argtypes = [I, I, P, L, P, L, P, P]
    E.lx_emu_adj_gemm.argtypes = [I, I, I, I, I, D, P, L, P, L, D, P, L, I, I, I, P]
    return E


@pytest.mark.parametrize("jb", [64, 37, 1])
def test_adjoint_diag_kernel(jb):
    """lx_adj_diag_kernel: mode 0 = the block inverse, mode 1 = the level-2
    Cholesky adjoint of the block (oracle chol_level2_adjoint) and P + P^T."""
    from oracle import adjoint as adj
    E = _emu_lib()
    rng = np.random.default_rng(jb)
    G = rng.standard_normal((jb, jb))
    Lb = np.asfortranarray(np.linalg.cholesky(G @ G.T / jb + np.eye(jb)))
    Db = np.asfortranarray(np.tril(rng.standard_normal((jb, jb))))
    Z = np.zeros((64, 64), order="F")
    S = np.full((64, 64), 7.0, order="F")  # entries past the block's order stay untouched
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    E.lx_emu_adj_diag(0, jb, p(Lb), jb, p(Db), jb, p(Z), p(S))
    assert np.abs(Z[:jb, :jb] - np.linalg.inv(Lb)).max() <= 1e-12 * np.abs(np.linalg.inv(Lb)).max()
    assert np.array_equal(Z[jb:, jb:], np.eye(64 - jb))
    ref = adj.chol_level2_adjoint(Lb, Db)
    out = Db.copy(order="F")
    E.lx_emu_adj_diag(1, jb, p(Lb), jb, p(out), jb, p(Z), p(S))
    assert np.abs(np.tril(out) - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(np.triu(out, 1), np.triu(Db, 1))
    assert np.abs(S[:jb, :jb] - (ref + ref.T)).max() <= 1e-12 * np.abs(ref).max()
    assert (S[jb:, :] == 7.0).all() and (S[:, jb:] == 7.0).all()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_adjoint_gemm_kernel(ta, tb):
    """lx_adj_gemm_kernel: C = alpha op(A) op(B) + beta C on ragged shapes, the
    lower-only mask, and split-K partials."""
    E = _emu_lib()
    rng = np.random.default_rng(10 * ta + tb)
    M, N, K = 70, 45, 37
    op = lambda X, t: X.T if t else X
    A = np.asfortranarray(rng.standard_normal((K, M) if ta else (M, K)))
    B = np.asfortranarray(rng.standard_normal((N, K) if tb else (K, N)))
    C0 = np.asfortranarray(rng.standard_normal((M, N)))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    C = C0.copy(order="F")
    E.lx_emu_adj_gemm(ta, tb, M, N, K, -1.5, p(A), A.shape[0], p(B), B.shape[0], 0.5, p(C), M, 0, 1, 48, None)
    ref = -1.5 * op(A, ta) @ op(B, tb) + 0.5 * C0
    assert np.abs(C - ref).max() <= 1e-12 * np.abs(ref).max()
    # lower-only: entries above the diagonal untouched
    C = C0.copy(order="F")
    E.lx_emu_adj_gemm(ta, tb, M, N, K, 1.0, p(A), A.shape[0], p(B), B.shape[0], 1.0, p(C), M, 1, 1, 48, None)
    ref = op(A, ta) @ op(B, tb) + C0
    low = np.tril(np.ones((M, N), dtype=bool))
    assert np.abs(C[low] - ref[low]).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(C[~low], C0[~low])
    # split-K: three chunks of 16 (the last ragged), partials summed in order
    part = np.zeros(3 * M * N)
    E.lx_emu_adj_gemm(ta, tb, M, N, K, 2.0, p(A), A.shape[0], p(B), B.shape[0], 0.0, p(C), M, 0, 3, 16, p(part))
    tot = part.reshape(3, N, M).sum(0).T
    ref = 2.0 * op(A, ta) @ op(B, tb)
    assert np.abs(tot - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("variant", [0, 1])
def test_plans_cut_units_inside_tiles(emu, variant):
    """The hypermatrix path needs every unit's operands inside one b x b tile:
    the factor plans of n = mt * b over tile sizes, both diagonal sub-tiles the
    library uses (FP32 128, FP64 256), both update precisions and GEMM widths."""
    E = _load_emu()
    E.lx_emu_plan_in_tiles.argtypes = [ctypes.c_int] * 6
    for b in (1, 37, 64, 96, 128, 200, 256, 384, 512, 1024):
        for mt in (1, 2, 5):
            for up, sub, wn in ((0, 128, 512), (1, 128, 512), (0, 256, 512), (1, 128, 1024), (0, 64, 256)):
                assert E.lx_emu_plan_in_tiles(variant, mt * b, b, up, sub, wn) == 1, (b, mt, up, sub, wn)


@pytest.mark.parametrize("dtype,variant,mt,b", [(0, 0, 3, 48), (0, 1, 3, 48), (1, 0, 2, 80), (0, 0, 1, 64)])
def test_factor_hypermatrix(emu, orc, dtype, variant, mt, b):
    """The tiled executor instantiation on a hypermatrix (separate tile
    buffers): each tile of the factor against the oracle's CHOL_HRL / CHOL_HLL
    on the same tiles and against LAPACK on the assembled matrix."""
    E = _load_emu()
    E.lx_emu_run_tiles.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_int)]
    n = mt * b
    dt = np.float64 if dtype == 0 else np.float32
    A0 = orc.genmat_spd(n, 11, diag_mode=1)
    tiles = orc.to_tiles(A0.astype(dt), b)
    ref = {k: v.copy(order="F") for k, v in tiles.items()}
    assert orc.chol_tiles(ref, mt, b, "hrl" if variant == 0 else "hll") == 0
    tab = orc.tile_table(tiles, mt)
    info = ctypes.c_int(-1)
    assert E.lx_emu_run_tiles(dtype, variant, mt, b, ctypes.cast(tab, ctypes.c_void_p), ctypes.byref(info)) == 0
    assert info.value == 0
    L = np.linalg.cholesky(A0)
    tol = 1e-12 if dtype == 0 else 2e-6
    got = np.tril(orc.from_tiles(tiles, mt, b))
    assert np.abs(got - L).max() <= tol * np.abs(L).max()
    assert np.abs(got - np.tril(orc.from_tiles(ref, mt, b))).max() <= tol * np.abs(L).max()
    # the strict upper part of each diagonal tile is left as it was
    for k in range(mt):
        assert np.array_equal(np.triu(tiles[(k, k)], 1), np.triu(A0[k * b:(k + 1) * b, k * b:(k + 1) * b].astype(dt), 1))


@pytest.mark.parametrize("ta,tb,uplo", [(0, 0, 0), (1, 0, 2), (0, 1, 1), (1, 1, 2)])
def test_gemm64_float_and_upper(emu, ta, tb, uplo):
    """The shared 64 x 64-tile GEMM kernel in FP32 (v_mfma_f32_16x16x4f32),
    every operand orientation, with the lower and upper masks of SYRK."""
    E = _load_emu()
    F, I, L, P = ctypes.c_float, ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p
    E.lx_emu_gemm64s.argtypes = [I, I, I, I, I, F, P, L, P, L, F, P, L, I]
    M, N, K = 70, 90, 37
    rng = np.random.default_rng(ta * 4 + tb * 2 + uplo)
    A = np.asfortranarray(rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32))
    B = np.asfortranarray(rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32))
    C0 = np.asfortranarray(rng.standard_normal((M + 3, N)).astype(np.float32))
    C = C0.copy(order="F")
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    E.lx_emu_gemm64s(ta, tb, M, N, K, 1.5, p(A), A.shape[0], p(B), B.shape[0], -0.5, p(C), C.shape[0], uplo)
    opA = (A.T if ta else A).astype(np.float64)
    opB = (B.T if tb else B).astype(np.float64)
    ref = 1.5 * opA @ opB - 0.5 * C0[:M].astype(np.float64)
    r, c = np.indices((M, N))
    mask = np.ones((M, N), bool) if uplo == 0 else (r >= c if uplo == 1 else r <= c)
    assert np.abs(C[:M][mask] - ref[mask]).max() <= 1e-5 * np.abs(ref).max()
    assert np.array_equal(C[:M][~mask], C0[:M][~mask])
    assert np.array_equal(C[M:], C0[M:])


def _emu_raw():
    E = _load_emu()
    E.lx_emu_run_stuck.restype = ctypes.c_int
    E.lx_emu_run_stuck.argtypes = [ctypes.c_int] * 4 + [ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong,
                                                        ctypes.POINTER(ctypes.c_int)]
    E.lx_emu_sweep_ex.restype = ctypes.c_int
    E.lx_emu_sweep_ex.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
                                                       ctypes.c_longlong, ctypes.c_int, ctypes.c_longlong]
    return E


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_executor_bounded_wait_aborts(orc, dtype):
    """A dependency that is never met ends through the executor's bounded wait:
    the wave that gave up sets the abort flag, counts one timeout, and every
    wave leaves the persistent loop (the call returns, all 256 emulated
    threads joined) with fewer units done than listed -- the condition the
    host turns into LINOMPSS_ERR_SCHED."""
    E = _emu_raw()
    n, b = 300, 64
    A = orc.genmat_spd(n, 3, diag_mode=1).astype(dtype, order="F")
    st = (ctypes.c_int * 4)()
    u = E.lx_emu_run_stuck(0 if dtype == np.float64 else 1, n, b, 5, 100000, A.ctypes.data_as(ctypes.c_void_p),
                           n, st)
    assert u >= 5
    abort, timeouts, done, nunits = list(st)
    assert abort == 1 and timeouts == 1
    assert done == u < nunits


@pytest.mark.parametrize("trans", [0, 1])
def test_sweep_bounded_wait_aborts(trans):
    """A step whose partial sums never arrive: its solve unit gives up after
    spin_ticks, sets the abort flag, and the remaining waits end at once, so
    the launch drains (rc -3 = the abort flag, which the host reports as
    LINOMPSS_ERR_SCHED)."""
    E = _emu_raw()
    n, nr = 700, 2
    rng = np.random.default_rng(9)
    L = np.asfortranarray(np.tril(rng.uniform(0, 1, (n, n))) + n * np.eye(n))
    X = np.asfortranarray(rng.uniform(-1, 1, (n, nr)))
    rc = E.lx_emu_sweep_ex(0, trans, n, nr, L.ctypes.data_as(ctypes.c_void_p), n, X.ctypes.data_as(ctypes.c_void_p),
                           n, 6, 100000)
    assert rc == -3


@pytest.mark.parametrize("dtype,bits", [(np.float64, np.uint64), (np.float32, np.uint32)])
@pytest.mark.parametrize("trans", [0, 1])
def test_sweep_nan_with_mailbox_bits(orc, dtype, bits, trans):
    """An input NaN whose bits equal the mailbox's 'empty' mark (all ones)
    propagates as NaN: the solve unit publishes the canonical quiet NaN in its
    place, so no reader waits for a value that looks unwritten (rc 0, no
    abort)."""
    E = _emu_raw()
    n, nr, row = 300, 2, 70
    rng = np.random.default_rng(10)
    L = np.asfortranarray(np.tril(rng.uniform(0.5, 1, (n, n))) + n * np.eye(n)).astype(dtype, order="F")
    X = np.asfortranarray(rng.uniform(-1, 1, (n, nr))).astype(dtype, order="F")
    X.view(bits)[row, 0] = np.iinfo(bits).max
    assert np.isnan(X[row, 0])
    rc = E.lx_emu_sweep_ex(0 if dtype == np.float64 else 1, trans, n, nr, L.ctypes.data_as(ctypes.c_void_p), n,
                           X.ctypes.data_as(ctypes.c_void_p), n, -1, 100000000 * 20)
    assert rc == 0
    # forward: rows from `row` on depend on x[row]; backward: rows up to `row`.
    # The solve multiplies by the inverse of the 64 x 64 diagonal block, whose
    # zeros times NaN are NaN, so the whole 64-row block of `row` turns NaN
    blk = row // 64 * 64
    dep = np.arange(n) >= blk if trans == 0 else np.arange(n) < blk + 64
    assert np.isnan(X[dep, 0]).all()
    assert np.isfinite(X[~dep, 0]).all()
    assert np.isfinite(X[:, 1]).all()
