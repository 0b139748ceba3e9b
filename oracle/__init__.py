"""CPU oracle of the tiled Cholesky path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker or the timed CPU
baseline.  The product (``linompss_amd``) never imports it.

It loads ``oracle/_build/liblx_oracle.so`` (built by ``oracle/Makefile``, a
plain-C restatement of the reference's loop nests over host LAPACK/BLAS; see
``oracle/lx_oracle.h`` for the reference file:line of every function).

Parity pins (each function is checked by ``tests/test_oracle.py``):
* ``drand48_fill``   -- bit-exact against glibc ``drand48`` (POSIX definition)
* ``genmat_spd`` etc -- fill order against the reference's loop nest, restated
  in numpy from the same drand48 stream
* ``dchol_rl/ll``, ``schol_rl/ll`` -- against numpy/LAPACK ``cholesky``
  (uniqueness of the factor), the reference's ``potrf_test`` residual bound,
  and a brute-force triple loop on tiny sizes
* ``dltrsm_lower``   -- against scipy ``solve_triangular`` and brute force
* ``dspdsolv``       -- backward error and ``numpy.linalg.solve``
* ``dposv_ir``       -- backward-error target, fallback trigger on a matrix
  that the FP32 factor cannot refine
* ``relerr_lower``, ``potrf_resid``, ``backward_error`` -- closed forms in
  numpy on random inputs and hand-made cases
* ``genmat_covariant``, ``genmat_costas``, ``dgenmat_config`` -- exact entries
  fixed by the definition (integer powers, square roots of perfect squares),
  the Toeplitz structure, the SPD bound of the Toeplitz symbol, and the
  drand48 stream
* ``philox4x32_10`` -- Random123's published known-answer vectors;
  ``gen_uniform`` / ``gen_gauss`` -- the stream's moments and the Box-Muller
  identity g0^2 + g1^2 = -2 ln u0
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "_build" / "liblx_oracle.so"

os.environ.setdefault("OMP_MAX_TASK_PRIORITY", "10000")


def build() -> Path:
    import subprocess
    subprocess.run(["make", "-s", "-C", str(_DIR)], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists():
        build()
    # MKL loads its computational kernels with dlopen; they resolve symbols
    # from libmkl_core, so the oracle is loaded with RTLD_GLOBAL.
    return ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)


lib = _load()
_i, _d, _f, _p, _l, _i64 = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p, ctypes.c_long, ctypes.c_int64


class IrOpts(ctypes.Structure):
    _fields_ = [("max_sweeps", _i), ("tol", _d), ("stall_ratio", _d), ("nthreads", _i)]


class IrStatus(ctypes.Structure):
    _fields_ = [("sweeps", _i), ("fallback", _i), ("info", _i), ("berr", _d)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("lxo_drand48_fill", None, _l, _i64, _p)
_sig("lxo_genmat_spd", None, _i, _i, _l, _i, _p)
_sig("lxo_genmat", None, _i, _i, _i, _l, _p)
_sig("lxo_genmat_sym_full", None, _i, _l, _p)
for nm in ("lxo_dchol_rl", "lxo_dchol_ll", "lxo_schol_rl", "lxo_schol_ll"):
    _sig(nm, _i, _i, _i, _p, _i, _i)
for nm in ("lxo_dchol_hrl", "lxo_dchol_hll", "lxo_schol_hrl", "lxo_schol_hll"):
    _sig(nm, _i, _i, _i, _p, _i)
_sig("lxo_dltrsm_lower", None, _i, _i, _i, _i, _i, _d, _p, _i, _p, _i, _i)
_sig("lxo_sltrsm_lower", None, _i, _i, _i, _i, _i, _f, _p, _i, _p, _i, _i)
_sig("lxo_dspdsolv", _i, _i, _i, _i, _p, _p, _i)
_sig("lxo_dposv_ir", _i, _i, _i, _i, _p, _i, _p, _i, _p, _i, ctypes.POINTER(IrOpts), ctypes.POINTER(IrStatus))
_sig("lxo_relerr_lower", _d, _i, _p, _i, _p, _i)
_sig("lxo_potrf_resid", _d, _i, _p, _i, _p, _i)
_sig("lxo_backward_error", _d, _i, _i, _p, _i, _p, _i, _p, _i)
_sig("lxo_genmat_covariant", _i, _i, _d, _d, _p)
_sig("lxo_genmat_costas", None, _i, _i, _l, _p)
_sig("lxo_dgenmat_config", _i, _i, _i, _i, _i, _p, _l, _p)
_sig("lxo_philox4x32_10", None, _p, _p, _p)
_sig("lxo_gen_uniform", None, _i, _i, ctypes.c_uint32, _d, _d, _p, _i)
_sig("lxo_gen_gauss", None, _i, _i, ctypes.c_uint32, _p, _i)
_sig("lxo_wtime", _d)
_sig("lxo_max_threads", _i)


def _pt(a):
    return a.ctypes.data_as(_p)


def drand48_fill(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib.lxo_drand48_fill(seed, count, _pt(out))
    return out


def genmat_spd(n: int, seed: int, diag_mode: int = 1, lda: int | None = None) -> np.ndarray:
    lda = n if lda is None else lda
    A = np.zeros((lda, n), dtype=np.float64, order="F")
    lib.lxo_genmat_spd(n, lda, seed, diag_mode, _pt(A))
    return A


def genmat(m: int, n: int, seed: int) -> np.ndarray:
    A = np.zeros((m, n), dtype=np.float64, order="F")
    lib.lxo_genmat(m, n, m, seed, _pt(A))
    return A


def genmat_sym_full(n: int, seed: int) -> np.ndarray:
    A = np.zeros((n, n), dtype=np.float64, order="F")
    lib.lxo_genmat_sym_full(n, seed, _pt(A))
    return A


def genmat_covariant(n: int, p: float = 2.0, k: float = 2.0) -> np.ndarray | None:
    """GENMAT_COVARIANT (genmat.c:263-302); None for p <= 0 (the reference's NULL)."""
    A = np.zeros((n, n), dtype=np.float64, order="F")
    return A if lib.lxo_genmat_covariant(n, p, k, _pt(A)) == 0 else None


def genmat_costas(m: int, n: int, seed: int) -> np.ndarray:
    A = np.zeros((m, n), dtype=np.float64, order="F")
    lib.lxo_genmat_costas(m, n, seed, _pt(A))
    return A


# ompssmat_t (src/matutil/ompss_mat.h:5-14)
GENMATCONF = {"spd": 0, "covar": 1, "rand": 2, "eye": 3, "diagdom": 4, "ones": 5, "zero": 6, "costas": 7}


def dgenmat_config(kind, m: int, n: int, ldim: int, pars=None, seed: int = 0) -> np.ndarray | None:
    """dgenmat_config (genmat_config.c:34-133) with explicit parameters and seed;
    None where the reference returns NULL.  The array is ldim x n (ld m for
    COVAR and COSTAS, which the reference allocates at m x m / m x n)."""
    k = GENMATCONF[kind] if isinstance(kind, str) else int(kind)
    rows = m if k in (1, 7) else ldim
    cols = m if k in (0, 1, 3, 6) else n
    A = np.zeros((rows, cols), dtype=np.float64, order="F")
    pa = None if pars is None else (ctypes.c_double * 4)(*(list(pars) + [2.0] * (4 - len(pars))))
    rc = lib.lxo_dgenmat_config(k, m, n, ldim, pa, seed, _pt(A))
    return A if rc == 0 else None


def philox4x32_10(ctr, key) -> tuple:
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib.lxo_philox4x32_10(c, k, o)
    return tuple(o)


def gen_uniform(m: int, n: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    X = np.zeros((m, n), dtype=np.float64, order="F")
    lib.lxo_gen_uniform(m, n, seed, lo, hi, _pt(X), m)
    return X


def gen_gauss(m: int, n: int, seed: int) -> np.ndarray:
    G = np.zeros((m, n), dtype=np.float64, order="F")
    lib.lxo_gen_gauss(m, n, seed, _pt(G), m)
    return G


def chol(A: np.ndarray, b: int, variant: str = "rl", nthreads: int = 0) -> int:
    """Tiled Cholesky in place on a column-major array (float64 or float32)."""
    assert A.flags.f_contiguous
    pre = "lxo_d" if A.dtype == np.float64 else "lxo_s"
    return getattr(lib, f"{pre}chol_{variant}")(A.shape[1], b, _pt(A), A.shape[0], nthreads)


def to_tiles(A: np.ndarray, b: int) -> dict:
    """Hypermatrix form of the lower triangle of A (n a multiple of b): tile
    (i, j), i >= j, as a contiguous column-major b x b array (chol_hsetup.c)."""
    n = A.shape[0]
    assert n % b == 0
    mt = n // b
    return {(i, j): np.array(A[i * b:(i + 1) * b, j * b:(j + 1) * b], order="F")
            for j in range(mt) for i in range(j, mt)}


def from_tiles(tiles: dict, mt: int, b: int) -> np.ndarray:
    """Column-major matrix with the tiles' values in the lower tiles, zeros above."""
    dt = next(iter(tiles.values())).dtype
    A = np.zeros((mt * b, mt * b), dtype=dt, order="F")
    for (i, j), t in tiles.items():
        A[i * b:(i + 1) * b, j * b:(j + 1) * b] = t
    return A


def tile_table(tiles: dict, mt: int):
    """ctypes array of mt*mt pointers, entry j*mt + i = tile (i, j) (null above)."""
    tab = (ctypes.c_void_p * (mt * mt))()
    for (i, j), t in tiles.items():
        tab[j * mt + i] = t.ctypes.data
    return tab


def chol_tiles(tiles: dict, mt: int, b: int, variant: str = "hrl", nthreads: int = 0) -> int:
    """CHOL_HRL / CHOL_HLL in place on the tiles."""
    pre = "lxo_d" if next(iter(tiles.values())).dtype == np.float64 else "lxo_s"
    tab = tile_table(tiles, mt)
    return getattr(lib, f"{pre}chol_{variant}")(mt, b, ctypes.cast(tab, _p), nthreads)


def ltrsm_lower(trans: int, A: np.ndarray, B: np.ndarray, bm: int, alpha: float = 1.0, nthreads: int = 0) -> None:
    assert A.flags.f_contiguous and B.flags.f_contiguous
    bn = B.shape[1] if B.ndim == 2 else 1
    fn = lib.lxo_dltrsm_lower if A.dtype == np.float64 else lib.lxo_sltrsm_lower
    fn(trans, bm, bn, A.shape[0], bn, alpha, _pt(A), A.shape[0], _pt(B), B.shape[0], nthreads)


def dspdsolv(A: np.ndarray, B: np.ndarray, b: int, nthreads: int = 0) -> int:
    m = A.shape[0]
    nr = B.shape[1] if B.ndim == 2 else 1
    return lib.lxo_dspdsolv(m, nr, b, _pt(A), _pt(B), nthreads)


def dposv_ir(A, B, b, max_sweeps=30, tol=1e-14, stall_ratio=0.5, nthreads=0):
    B = np.asfortranarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nr = A.shape[0], B.shape[1]
    X = np.zeros((n, nr), order="F")
    st = IrStatus()
    o = IrOpts(max_sweeps, tol, stall_ratio, nthreads)
    info = lib.lxo_dposv_ir(n, nr, b, _pt(A), A.shape[0], _pt(B), n, _pt(X), n, ctypes.byref(o), ctypes.byref(st))
    return X, st, info


def relerr_lower(A, L) -> float:
    return lib.lxo_relerr_lower(A.shape[0], _pt(A), A.shape[0], _pt(L), L.shape[0])


def potrf_resid(A0, L) -> float:
    return lib.lxo_potrf_resid(A0.shape[0], _pt(A0), A0.shape[0], _pt(L), L.shape[0])


def backward_error(A, X, B) -> float:
    X = np.asfortranarray(X.reshape(X.shape[0], -1), dtype=np.float64)
    B = np.asfortranarray(B.reshape(B.shape[0], -1), dtype=np.float64)
    return lib.lxo_backward_error(A.shape[0], X.shape[1], _pt(A), A.shape[0], _pt(X), X.shape[0], _pt(B),
                                  B.shape[0])


def wtime() -> float:
    return lib.lxo_wtime()


def max_threads() -> int:
    return lib.lxo_max_threads()
