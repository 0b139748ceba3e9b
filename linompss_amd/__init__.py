"""linompss_amd -- MI355X (gfx950) tiled Cholesky / SPD solve / refinement.

Python host mirror of Linompss's C entry points for this path, as ctypes
calls into ``liblinompss.so`` (built in-tree by :mod:`linompss_amd.build`).
The names and argument meanings follow the reference's interface:

* ``ompss_dchol_rl(n, b, t, A, lda)`` and its ``ll`` / ``s`` twins
  (src/cholesky/dense/ompss_dense_chol.h:32-38),
* the hypermatrix forms ``ompss_{d,s}chol_{hrl,hll,nhrl,nhll}(mt, b, t, Ah)``
  (ompss_dense_chol.h:15-27,41-47) over a dict of b x b tiles,
* ``dspdsolv_main(m, n, b, A, B)`` (src/spdsolv/spdsolv_main.h:20),
* ``bblas_dltrsm`` lower branches (src/blocked_blas/bblas_trsm.h:116-160),
* an iterative-refinement solve keeping itref's loop contract
  (src/itref/itref.c:999-1101), for one system or for a batch of many
  independent systems of uneven order in one persistent launch.

Host-array functions take column-major (Fortran-ordered) numpy arrays and
work in place like the reference.  Device functions take :class:`DeviceMatrix`
objects (column-major device buffers allocated through the library).

There is no CPU fallback: if the library is missing or cannot be loaded,
importing this package raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

__all__ = [
    "lib", "LinompssError", "ompss_dchol_rl", "ompss_dchol_ll", "ompss_schol_rl", "ompss_schol_ll",
    "dspdsolv_main", "sspdsolv_main", "dltrsm", "sltrsm", "dposv_ir", "chol_dev", "posv_dev",
    "potrs_dev", "dposv_ir_dev", "last_exec_stats", "plan_tdg", "plan_info", "ExecOpts", "IrOpts",
    "IrStatus", "ExecStats", "DeviceMatrix", "device_count", "BatchOpts", "BatchStatus", "DeviceBatch",
    "BatchPlan", "dposv_ir_batched_dev", "gen_batch_spd", "ompss_dchol_hrl", "ompss_dchol_hll",
    "ompss_schol_hrl", "ompss_schol_hll", "ompss_dchol_nhrl", "ompss_dchol_nhll", "ompss_schol_nhrl",
    "ompss_schol_nhll", "chol_tiles_dev", "gemm", "syrk", "trsm", "write_dense", "read_dense", "read_values",
    "blas_desc", "lapack_desc", "exported_desc", "CBLAS",
]

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("LINOMPSS_LIB", _PKG / "liblinompss.so"))

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python linompss_amd/build.py` (hipcc, gfx950). "
        "There is no CPU fallback.")
# The library binds to the HIP runtime of the ROCm it was built with
# (/opt/rocm).  PyTorch bundles an older libamdhip64 with the same soname; a
# process that initialises torch's runtime first would make this library run
# on that one, a pairing these kernels are not validated with, so the package
# does not import torch and its device API works on raw device pointers.
lib = ctypes.CDLL(str(LIB_PATH))

_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_f = ctypes.c_float
_p = ctypes.c_void_p
_c = ctypes.c_char
_ip = ctypes.POINTER(ctypes.c_int)


class ExecOpts(ctypes.Structure):
    _fields_ = [("policy", _i), ("unit_m", _i), ("unit_n", _i), ("wg_per_cu", _i), ("timeout_s", _d),
                ("fused_solve", _i), ("update_prec", _i)]


class ExecStats(ctypes.Structure):
    _fields_ = [("exec_ms", _d), ("units", _i), ("tasks", _i), ("grid", _i), ("sim_ms", _d)]


class IrOpts(ctypes.Structure):
    _fields_ = [("max_sweeps", _i), ("tol", _d), ("stall_ratio", _d), ("update_prec", _i)]


class IrStatus(ctypes.Structure):
    _fields_ = [("sweeps", _i), ("fallback", _i), ("info", _i), ("berr", _d), ("t_factor_s", _d),
                ("t_total_s", _d)]


def _sig(name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_sig("ompss_dchol_rl", _i, _i, _i, _i, _p, _i)
_sig("ompss_dchol_ll", _i, _i, _i, _i, _p, _i)
_sig("ompss_schol_rl", _i, _i, _i, _i, _p, _i)
_sig("ompss_schol_ll", _i, _i, _i, _i, _p, _i)
for _nm in ("ompss_dchol_hrl", "ompss_dchol_hll", "ompss_schol_hrl", "ompss_schol_hll", "ompss_dchol_nhrl",
            "ompss_dchol_nhll", "ompss_schol_nhrl", "ompss_schol_nhll"):
    _sig(_nm, _i, _i, _i, _i, _p)
_sig("dspdsolv_main", _i, _i, _i, _i, _p, _p)
_sig("sspdsolv_main", _i, _i, _i, _i, _p, _p)
_sig("linompss_dltrsm", _i, _c, _c, _c, _i, _i, _i, _i, _d, _p, _i, _p, _i)
_sig("linompss_sltrsm", _i, _c, _c, _c, _i, _i, _i, _i, _f, _p, _i, _p, _i)
_sig("linompss_dposv_ir", _i, _i, _i, _i, _p, _i, _p, _i, _p, _i, ctypes.POINTER(IrOpts),
     ctypes.POINTER(IrStatus))
_sig("linompss_chol_dev", _i, _i, _i, _i, _i, _p, _i64, _p, ctypes.POINTER(ExecOpts))
_sig("linompss_chol_tiles_dev", _i, _i, _i, _i, _i, _p, _p, ctypes.POINTER(ExecOpts))
_sig("linompss_dgemm", _i, _c, _c, _i, _i, _i, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i)
_sig("linompss_sgemm", _i, _c, _c, _i, _i, _i, _i, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i)
_sig("linompss_dsyrk", _i, _c, _c, _i, _i, _i, _d, _p, _i, _d, _p, _i)
_sig("linompss_ssyrk", _i, _c, _c, _i, _i, _i, _f, _p, _i, _f, _p, _i)
_sig("linompss_dtrsm", _i, _c, _c, _c, _c, _i, _i, _i, _d, _p, _i, _p, _i)
_sig("linompss_strsm", _i, _c, _c, _c, _c, _i, _i, _i, _f, _p, _i, _p, _i)
_sig("linompss_fprint_ddense2mm", _i, ctypes.c_char_p, ctypes.c_char_p, _i, _i, _p, _i)
_sig("linompss_fprint_sdense2mm", _i, ctypes.c_char_p, ctypes.c_char_p, _i, _i, _p, _i)
_sig("linompss_read_mm2ddense", _i64, ctypes.c_char_p, _i64, _p)
_sig("linompss_read_mm2sdense", _i64, ctypes.c_char_p, _i64, _p)
_sig("linompss_dense2mm_dims", _i, ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i))
_sig("linompss_read_ddense2mm", _i, ctypes.c_char_p, _i, _i, _p, _i)
_sig("linompss_read_sdense2mm", _i, ctypes.c_char_p, _i, _i, _p, _i)
_sig("linompss_gemm_dev", _i, _i, _c, _c, _i, _i, _i, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p)
_sig("linompss_syrk_dev", _i, _i, _c, _c, _i, _i, _d, _p, _i64, _d, _p, _i64, _p)
_sig("linompss_trsm_blas_dev", _i, _i, _c, _c, _c, _c, _i, _i, _d, _p, _i64, _p, _i64, _p)
_sig("linompss_posv_dev", _i, _i, _i, _i, _i, _p, _i64, _p, _i64, _p, ctypes.POINTER(ExecOpts))
_sig("linompss_potrs_dev", _i, _i, _i, _i, _i, _p, _i64, _p, _i64, _p, ctypes.POINTER(ExecOpts))
_sig("linompss_potrs_dev_noalloc", _i, _i, _i, _i, _i, _p, _i64, _p, _i64, _p, ctypes.POINTER(ExecOpts))
_sig("linompss_dposv_ir_dev", _i, _i, _i, _i, _p, _i64, _p, _i64, _p, _i64, _p, ctypes.POINTER(IrOpts),
     ctypes.POINTER(IrStatus))
_sig("linompss_last_exec_stats", _i, ctypes.POINTER(ExecStats))
_sig("linompss_plan_tdg", _i, _i, _i, _i, _i, _ip, _ip, _i, _ip, _i, _ip, _ip)
_sig("linompss_plan_info", _i, _i, _i, _i, _i, ctypes.POINTER(ExecOpts), _ip, _ip, ctypes.POINTER(_d))
_sig("linompss_version", ctypes.c_char_p)

# The reference's flag descriptors ompssblas_t / ompsslapack_t
# (src/fp/fpmatr.h:14-35) are 8-byte unions of an int (a CBLAS value, its MKL
# build) or a char (LAPACK) with a char * ("N", "T", "U", "L", "R", its
# Fortran-BLAS build), passed by value.  The x86-64 calling convention passes
# such a union in one integer register, so ctypes carries it as the 64-bit
# integer: the CBLAS value / character, or the address of the string.
_u64 = ctypes.c_uint64
CBLAS = {"trans": {"N": 111, "T": 112}, "uplo": {"U": 121, "L": 122}, "diag": {"N": 131, "U": 132},
         "side": {"L": 141, "R": 142}}
_FLAG_STR = {c: ctypes.create_string_buffer(c.encode()) for c in "NTULRnturlx"}


def blas_desc(kind: str, letter: str, flavour: str = "enum") -> int:
    """An ompssblas_t as the integer ctypes passes: ``flavour`` "enum" gives
    the CBLAS value of ``letter`` for ``kind`` (trans, uplo, diag, side),
    "str" the address of a NUL-terminated one-letter string."""
    if flavour == "enum":
        return CBLAS[kind][letter.upper()]
    return ctypes.addressof(_FLAG_STR[letter])


def lapack_desc(letter: str, flavour: str = "enum") -> int:
    """An ompsslapack_t: the character itself ("enum") or the address of a string ("str")."""
    return ord(letter) if flavour == "enum" else ctypes.addressof(_FLAG_STR[letter])


def exported_desc(name: str) -> int:
    """The library's own descriptor instance ``name`` (OMPSSBLAS_TRANSP, OMPSSLAPACK_LOWERTRIANG, ...)."""
    return _u64.in_dll(lib, name).value


_sig("ompss_dgemm", _i, _u64, _u64, _i, _i, _i, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i)
_sig("ompss_sgemm", _i, _u64, _u64, _i, _i, _i, _i, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i)
_sig("ompss_dsyrk", None, _u64, _u64, _i, _i, _i, _d, _p, _i, _d, _p, _i)
_sig("ompss_ssyrk", None, _u64, _u64, _i, _i, _i, _f, _p, _i, _f, _p, _i)
_sig("ompss_dtrsm", None, _u64, _u64, _u64, _u64, _i, _i, _i, _d, _p, _i, _p, _i)
_sig("ompss_strsm", None, _u64, _u64, _u64, _u64, _i, _i, _i, _f, _p, _i, _p, _i)
_sig("task_dpotrf", None, _u64, _i, _p, _i, _i)
_sig("task_spotrf", None, _u64, _i, _p, _i, _i)
_sig("task_dtrsm", None, _u64, _u64, _u64, _u64, _i, _i, _d, _p, _i, _p, _i, _i)
_sig("task_strsm", None, _u64, _u64, _u64, _u64, _i, _i, _f, _p, _i, _p, _i, _i)
_sig("task_dsyrk", None, _u64, _u64, _i, _i, _d, _p, _i, _d, _p, _i, _i)
_sig("task_ssyrk", None, _u64, _u64, _i, _i, _f, _p, _i, _f, _p, _i, _i)
_sig("task_dgemm", None, _u64, _u64, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i, _i)
_sig("task_sgemm", None, _u64, _u64, _i, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i, _i)

ERR_BASE = -1000
ERR_NAMES = {-1001: "no usable HIP device", -1002: "HIP call failed", -1003: "device allocation failed",
             -1004: "device scheduler timed out", -1005: "planner rejected the shape",
             -1006: "matrix file could not be opened, written or parsed",
             -1007: "size beyond the fixed limit of a scratch-free entry point"}


class LinompssError(RuntimeError):
    def __init__(self, fn: str, code: int):
        what = ERR_NAMES.get(code, "invalid argument %d" % -code if code < 0 else "info %d" % code)
        super().__init__(f"{fn} returned {code} ({what})")
        self.code = code


def _check(fn: str, rc: int, allow_info: bool = True) -> int:
    """Raises on argument/runtime errors; returns LAPACK-style info (>= 0)."""
    if rc < 0 or (rc > 0 and not allow_info):
        raise LinompssError(fn, rc)
    return rc


def _colmajor(a: np.ndarray, dtype) -> None:
    if a.dtype != dtype:
        raise TypeError(f"expected {np.dtype(dtype).name}, got {a.dtype}")
    if not a.flags.f_contiguous:
        raise ValueError("expected a column-major (Fortran-ordered) array")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p)


# ---- host-memory drop-in entry points -------------------------------------

def _chol(name, dtype, A: np.ndarray, b: int, t: int = 0, n: int | None = None, lda: int | None = None) -> int:
    _colmajor(A, dtype)
    n = A.shape[1] if n is None else n
    lda = A.shape[0] if lda is None else lda
    return _check(name, getattr(lib, name)(n, b, t, _ptr(A), lda))


def ompss_dchol_rl(A: np.ndarray, b: int, t: int = 0, n=None, lda=None) -> int:
    """Right-looking tiled Cholesky of the lower triangle, in place (chol_rlmain.c:22-53)."""
    return _chol("ompss_dchol_rl", np.float64, A, b, t, n, lda)


def ompss_dchol_ll(A: np.ndarray, b: int, t: int = 0, n=None, lda=None) -> int:
    """Left-looking tiled Cholesky (chol_llmain.c:23-57)."""
    return _chol("ompss_dchol_ll", np.float64, A, b, t, n, lda)


def ompss_schol_rl(A: np.ndarray, b: int, t: int = 0, n=None, lda=None) -> int:
    return _chol("ompss_schol_rl", np.float32, A, b, t, n, lda)


def ompss_schol_ll(A: np.ndarray, b: int, t: int = 0, n=None, lda=None) -> int:
    return _chol("ompss_schol_ll", np.float32, A, b, t, n, lda)


def _tile_table(tiles: dict, mt: int, b: int, dtype):
    """ctypes table of mt*mt tile addresses, entry j*mt + i = tile (i, j);
    every lower tile must be present, b x b, column-major, of dtype."""
    tab = (ctypes.c_void_p * max(1, mt * mt))()
    for j in range(mt):
        for i in range(j, mt):
            a = tiles.get((i, j))
            if a is None:
                raise ValueError(f"tile ({i}, {j}) missing")
            _colmajor(a, dtype)
            if a.shape != (b, b):
                raise ValueError(f"tile ({i}, {j}) is {a.shape}, expected ({b}, {b})")
            tab[j * mt + i] = a.ctypes.data
    return tab


def _chol_tiles(name, dtype, tiles: dict, mt: int, b: int, t: int = 0) -> int:
    tab = _tile_table(tiles, mt, b, dtype) if mt > 0 else None
    return _check(name, getattr(lib, name)(mt, b, t, ctypes.cast(tab, _p) if tab is not None else None))


def ompss_dchol_hrl(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    """Right-looking Cholesky of a hypermatrix in place (chol_hrlmain.c:16-37):
    tiles[(i, j)], i >= j, are b x b column-major float64 arrays."""
    return _chol_tiles("ompss_dchol_hrl", np.float64, tiles, mt, b, t)


def ompss_dchol_hll(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    """Left-looking Cholesky of a hypermatrix in place (chol_hllmain.c:17-45)."""
    return _chol_tiles("ompss_dchol_hll", np.float64, tiles, mt, b, t)


def ompss_schol_hrl(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    return _chol_tiles("ompss_schol_hrl", np.float32, tiles, mt, b, t)


def ompss_schol_hll(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    return _chol_tiles("ompss_schol_hll", np.float32, tiles, mt, b, t)


def ompss_dchol_nhrl(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    """The nested variants (ompss_dense_chol.c:48-62): the same factor."""
    return _chol_tiles("ompss_dchol_nhrl", np.float64, tiles, mt, b, t)


def ompss_dchol_nhll(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    return _chol_tiles("ompss_dchol_nhll", np.float64, tiles, mt, b, t)


def ompss_schol_nhrl(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    return _chol_tiles("ompss_schol_nhrl", np.float32, tiles, mt, b, t)


def ompss_schol_nhll(tiles: dict, mt: int, b: int, t: int = 0) -> int:
    return _chol_tiles("ompss_schol_nhll", np.float32, tiles, mt, b, t)


def _blas_pre(a: np.ndarray) -> str:
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "s"
    raise TypeError(f"expected float64 or float32, got {a.dtype}")


def gemm(transa: str, transb: str, m: int, n: int, k: int, alpha: float, A: np.ndarray, B: np.ndarray,
         beta: float, C: np.ndarray, b: int = 128, d: int = 128, c: int = 128) -> None:
    """C[:m, :n] = alpha op(A) op(B) + beta C in place (ompss_{d,s}gemm, src/gemm/ompss_mm.h:23-27);
    leading dimensions are the arrays' row counts."""
    pre = _blas_pre(C)
    for a in (A, B, C):
        _colmajor(a, C.dtype)
    _check(f"linompss_{pre}gemm", getattr(lib, f"linompss_{pre}gemm")(
        transa.encode(), transb.encode(), b, d, c, m, n, k, alpha, _ptr(A), A.shape[0], _ptr(B), B.shape[0],
        beta, _ptr(C), C.shape[0]), False)


def syrk(uplo: str, trans: str, n: int, k: int, alpha: float, A: np.ndarray, beta: float, C: np.ndarray,
         b: int = 128) -> None:
    """The uplo triangle of C[:n, :n] = alpha op(A) op(A)^T + beta C (ompss_{d,s}syrk, syrk_main.c)."""
    pre = _blas_pre(C)
    for a in (A, C):
        _colmajor(a, C.dtype)
    _check(f"linompss_{pre}syrk", getattr(lib, f"linompss_{pre}syrk")(
        uplo.encode(), trans.encode(), b, n, k, alpha, _ptr(A), A.shape[0], beta, _ptr(C), C.shape[0]), False)


def trsm(side: str, uplo: str, trans: str, diag: str, m: int, n: int, alpha: float, A: np.ndarray,
         B: np.ndarray, b: int = 128) -> None:
    """B[:m, :n] := X with op(A) X = alpha B (side L) or X op(A) = alpha B (side R)
    (ompss_{d,s}trsm, trsm_main.c:6-17 -> bblas_trsm.h:116-201)."""
    pre = _blas_pre(B)
    for a in (A, B):
        _colmajor(a, B.dtype)
    _check(f"linompss_{pre}trsm", getattr(lib, f"linompss_{pre}trsm")(
        side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, b, alpha, _ptr(A), A.shape[0],
        _ptr(B), B.shape[0]), False)


def write_dense(fname, A: np.ndarray, name: str = "A") -> None:
    """A (2-D, float64 or float32) to fname in the Octave text format of
    print_dense2mm (src/matutil/matfprint.c:104-124)."""
    A = np.asfortranarray(A)
    pre = _blas_pre(A)
    m, n = A.shape
    _check(f"linompss_fprint_{pre}dense2mm", getattr(lib, f"linompss_fprint_{pre}dense2mm")(
        os.fsencode(fname), name.encode(), m, n, _ptr(A), max(1, m)), False)


def read_dense(fname, dtype=np.float64) -> np.ndarray:
    """The matrix of an Octave text file, column-major, in the layout it was written."""
    m, n = _i(), _i()
    _check("linompss_dense2mm_dims", lib.linompss_dense2mm_dims(os.fsencode(fname), ctypes.byref(m), ctypes.byref(n)),
           False)
    A = np.zeros((m.value, n.value), dtype=dtype, order="F")
    pre = _blas_pre(A)
    _check(f"linompss_read_{pre}dense2mm", getattr(lib, f"linompss_read_{pre}dense2mm")(
        os.fsencode(fname), m.value, n.value, _ptr(A), max(1, m.value)), False)
    return A


def read_values(fname, count: int, dtype=np.float64) -> np.ndarray:
    """Up to count values in file order (read_mm2{d,s}dense, src/matutil/matfread.c:13-27)."""
    out = np.zeros(max(1, count), dtype=dtype)
    pre = _blas_pre(out)
    got = getattr(lib, f"linompss_read_mm2{pre}dense")(os.fsencode(fname), count, _ptr(out))
    _check(f"linompss_read_mm2{pre}dense", int(got) if got < 0 else 0, False)
    return out[:got]


def dspdsolv_main(A: np.ndarray, B: np.ndarray, b: int) -> int:
    """Factor A (m x m, lda = m) and solve A X = B in place (spdsolv_main.c:10-20)."""
    _colmajor(A, np.float64)
    _colmajor(B, np.float64)
    m = A.shape[0]
    if A.shape != (m, m) or B.shape[0] != m:
        raise ValueError("dspdsolv_main expects A m x m and B m x n with lda = ldb = m")
    nr = B.shape[1] if B.ndim == 2 else 1
    return _check("dspdsolv_main", lib.dspdsolv_main(m, nr, b, _ptr(A), _ptr(B)))


def sspdsolv_main(A: np.ndarray, B: np.ndarray, b: int) -> int:
    _colmajor(A, np.float32)
    _colmajor(B, np.float32)
    m = A.shape[0]
    nr = B.shape[1] if B.ndim == 2 else 1
    return _check("sspdsolv_main", lib.sspdsolv_main(m, nr, b, _ptr(A), _ptr(B)))


def dltrsm(trans: str, A: np.ndarray, B: np.ndarray, bm: int, alpha: float = 1.0) -> None:
    """B := alpha op(L)^{-1} B with L the lower triangle of A (bblas_trsm.h:116-160)."""
    _colmajor(A, np.float64)
    _colmajor(B, np.float64)
    m = A.shape[0]
    bn = B.shape[1] if B.ndim == 2 else 1
    _check("linompss_dltrsm", lib.linompss_dltrsm(b"L", trans.encode(), b"N", bm, bn, m, bn, alpha, _ptr(A),
                                                   A.shape[0], _ptr(B), B.shape[0]), allow_info=False)


def sltrsm(trans: str, A: np.ndarray, B: np.ndarray, bm: int, alpha: float = 1.0) -> None:
    _colmajor(A, np.float32)
    _colmajor(B, np.float32)
    m = A.shape[0]
    bn = B.shape[1] if B.ndim == 2 else 1
    _check("linompss_sltrsm", lib.linompss_sltrsm(b"L", trans.encode(), b"N", bm, bn, m, bn, alpha, _ptr(A),
                                                   A.shape[0], _ptr(B), B.shape[0]), allow_info=False)


def dposv_ir(A: np.ndarray, B: np.ndarray, b: int, opts: IrOpts | None = None):
    """Mixed-precision refinement solve; returns (X, IrStatus)."""
    _colmajor(A, np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nr = A.shape[0], B.shape[1]
    X = np.zeros((n, nr), dtype=np.float64, order="F")
    st = IrStatus()
    rc = lib.linompss_dposv_ir(n, nr, b, _ptr(A), A.shape[0], _ptr(B), n, _ptr(X), n,
                               ctypes.byref(opts) if opts is not None else None, ctypes.byref(st))
    _check("linompss_dposv_ir", rc)
    return X, st


# ---- device memory ------------------------------------------------------------

_sig("linompss_device_count", _i)
_sig("linompss_dev_alloc", _p, ctypes.c_size_t)
_sig("linompss_dev_free", _i, _p)
_sig("linompss_memcpy", _i, _p, _p, ctypes.c_size_t, _i)
_sig("linompss_dev_sync", _i)


def device_count() -> int:
    return int(lib.linompss_device_count())


class DeviceMatrix:
    """Column-major rows x cols matrix in device memory (leading dimension ld)."""

    def __init__(self, rows: int, cols: int, dtype=np.float64, ld: int | None = None):
        self.rows, self.cols = int(rows), int(cols)
        self.dtype = np.dtype(dtype)
        self.ld = max(1, int(ld if ld is not None else rows))
        self.nbytes = self.ld * max(1, self.cols) * self.dtype.itemsize
        p = lib.linompss_dev_alloc(self.nbytes)
        if not p:
            raise LinompssError("linompss_dev_alloc", -1003)
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceMatrix":
        a = np.asfortranarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1, order="F")
        m = cls(a.shape[0], a.shape[1], a.dtype)
        m.upload(a)
        return m

    def upload(self, a: np.ndarray) -> None:
        a = np.asfortranarray(a, dtype=self.dtype).reshape(self.rows, -1, order="F")
        if self.ld != self.rows:
            raise ValueError("upload needs ld == rows")
        _check("linompss_memcpy", lib.linompss_memcpy(self.ptr, a.ctypes.data_as(_p), a.nbytes, 1), False)

    def to_host(self) -> np.ndarray:
        out = np.empty((self.ld, self.cols), dtype=self.dtype, order="F")
        _check("linompss_memcpy", lib.linompss_memcpy(out.ctypes.data_as(_p), self.ptr, out.nbytes, 2), False)
        return out[: self.rows]

    def cols_to_host(self, c0: int, c1: int) -> np.ndarray:
        """Columns c0 .. c1-1 (all ld rows read, the first rows returned): a
        panel of a matrix too large to bring to the host at once."""
        if not (0 <= c0 <= c1 <= self.cols):
            raise ValueError(f"columns {c0}..{c1} outside 0..{self.cols}")
        out = np.empty((self.ld, c1 - c0), dtype=self.dtype, order="F")
        src = ctypes.c_void_p(self.ptr + c0 * self.ld * self.dtype.itemsize)
        _check("linompss_memcpy", lib.linompss_memcpy(out.ctypes.data_as(_p), src, out.nbytes, 2), False)
        return out[: self.rows]

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib.linompss_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dt(m: DeviceMatrix) -> int:
    if m.dtype == np.float64:
        return 0
    if m.dtype == np.float32:
        return 1
    raise TypeError(f"unsupported dtype {m.dtype}")


def _s(stream):
    return ctypes.c_void_p(stream) if stream else None


def chol_tiles_dev(tiles: dict, mt: int, b: int, variant: str = "rl", opts: ExecOpts | None = None,
                   stream=None) -> int:
    """Factor a hypermatrix held in device memory: tiles[(i, j)], i >= j, are
    b x b DeviceMatrix tiles (ld b) of one dtype; returns info."""
    ptrs = np.zeros(max(1, mt * mt), dtype=np.uint64)
    dts = set()
    for j in range(mt):
        for i in range(j, mt):
            t = tiles[(i, j)]
            if (t.rows, t.cols, t.ld) != (b, b, b):
                raise ValueError(f"tile ({i}, {j}) must be {b} x {b} with ld {b}")
            ptrs[j * mt + i] = t.ptr
            dts.add(t.dtype)
    if len(dts) > 1:
        raise TypeError("tiles of mixed dtypes")
    dt = 0 if dts == {np.dtype(np.float64)} else 1
    tab = DeviceMatrix(ptrs.size, 1)  # 8-byte entries hold the device addresses
    try:
        tab.upload(ptrs.view(np.float64))
        rc = lib.linompss_chol_tiles_dev(dt, 0 if variant == "rl" else 1, mt, b, tab.ptr, _s(stream),
                                         ctypes.byref(opts) if opts is not None else None)
    finally:
        tab.free()
    return _check("linompss_chol_tiles_dev", rc)


def chol_dev(A: DeviceMatrix, b: int, variant: str = "rl", opts: ExecOpts | None = None, stream=None) -> int:
    """Factor the lower triangle of the device matrix A in place; returns info."""
    rc = lib.linompss_chol_dev(_dt(A), 0 if variant == "rl" else 1, A.cols, b, A.ptr, A.ld, _s(stream),
                               ctypes.byref(opts) if opts is not None else None)
    return _check("linompss_chol_dev", rc)


def posv_dev(A: DeviceMatrix, X: DeviceMatrix, b: int, opts: ExecOpts | None = None, stream=None) -> int:
    """Factor A and solve A Y = X in place (one persistent launch); returns info."""
    rc = lib.linompss_posv_dev(_dt(A), A.cols, X.cols, b, A.ptr, A.ld, X.ptr, X.ld, _s(stream),
                               ctypes.byref(opts) if opts is not None else None)
    return _check("linompss_posv_dev", rc)


def potrs_dev(L: DeviceMatrix, X: DeviceMatrix, b: int, opts: ExecOpts | None = None, stream=None) -> None:
    rc = lib.linompss_potrs_dev(_dt(L), L.cols, X.cols, b, L.ptr, L.ld, X.ptr, X.ld, _s(stream),
                                ctypes.byref(opts) if opts is not None else None)
    _check("linompss_potrs_dev", rc, allow_info=False)


def potrs_dev_noalloc(L: DeviceMatrix, X: DeviceMatrix, b: int, stream=None) -> None:
    """potrs_dev on the library's fixed device array instead of allocated
    scratch (n <= 32768): for callers that cannot allocate."""
    rc = lib.linompss_potrs_dev_noalloc(_dt(L), L.cols, X.cols, b, L.ptr, L.ld, X.ptr, X.ld, _s(stream), None)
    _check("linompss_potrs_dev_noalloc", rc, allow_info=False)


def dposv_ir_dev(A: DeviceMatrix, B: DeviceMatrix, X: DeviceMatrix, b: int, opts: IrOpts | None = None,
                 stream=None) -> IrStatus:
    st = IrStatus()
    rc = lib.linompss_dposv_ir_dev(A.cols, B.cols, b, A.ptr, A.ld, B.ptr, B.ld, X.ptr, X.ld, _s(stream),
                                   ctypes.byref(opts) if opts is not None else None, ctypes.byref(st))
    _check("linompss_dposv_ir_dev", rc)
    return st


_sig("linompss_gen_gauss_spd", _i, _i, ctypes.c_uint32, _d, _p, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d))
_sig("linompss_gen_uniform", _i, _i, _i, _i, ctypes.c_uint32, _d, _d, _p, _i64)
_sig("linompss_gen_gauss", _i, _i, _i, ctypes.c_uint32, _p, _i64)
_sig("linompss_gen_covariant", _i, _i, _d, _d, _p, _i64)
_sig("linompss_gen_drand48_spd", _i, _i, _i, ctypes.c_long, _p, _i64)
_sig("linompss_gen_drand48", _i, _i, _i, _i, ctypes.c_long, _p, _i64)
_sig("linompss_relerr_lower_dev", _i, _i, _i, _p, _i64, _p, _i64, ctypes.POINTER(_d))


def gen_gauss_spd(n: int, seed: int, kappa: float) -> tuple[DeviceMatrix, float, float]:
    """A = G G^T / n + alpha I on the device with cond2(A) ~ kappa; returns (A, alpha, lambda_max)."""
    A = DeviceMatrix(n, n, np.float64)
    a, lm = _d(0.0), _d(0.0)
    _check("linompss_gen_gauss_spd", lib.linompss_gen_gauss_spd(n, seed, kappa, A.ptr, A.ld, ctypes.byref(a),
                                                                   ctypes.byref(lm)), False)
    return A, a.value, lm.value


def gen_uniform(m: int, n: int, seed: int, lo: float = -1.0, hi: float = 1.0, dtype=np.float64) -> DeviceMatrix:
    X = DeviceMatrix(m, n, dtype)
    _check("linompss_gen_uniform", lib.linompss_gen_uniform(0 if np.dtype(dtype) == np.float64 else 1, m, n, seed,
                                                             lo, hi, X.ptr, X.ld), False)
    return X


def gen_covariant(n: int, p: float = 2.0, k: float = 2.0, ld: int | None = None) -> DeviceMatrix:
    """The refinement app's default matrix GENMAT_COVARIANT (genmat.c:263-302)
    in device memory: |i - j|^-p off the diagonal, 1 + (j + 1)^(1/k) on it."""
    A = DeviceMatrix(n, n, ld=ld)
    _check("linompss_gen_covariant", lib.linompss_gen_covariant(n, p, k, A.ptr, A.ld), False)
    return A


def gen_drand48_spd(n: int, seed: int, dtype=np.float64, ld: int | None = None) -> DeviceMatrix:
    """The reference's GENMAT_SPD (genmat.c:49-70) in device memory, bit for
    bit: drand48 seeded by ``seed``, drawn column by column over i >= j,
    mirrored, n added to the diagonal."""
    A = DeviceMatrix(n, n, dtype, ld=ld)
    _check("linompss_gen_drand48_spd", lib.linompss_gen_drand48_spd(
        0 if np.dtype(dtype) == np.float64 else 1, n, seed, A.ptr, A.ld), False)
    return A


def gen_drand48(m: int, n: int, seed: int, dtype=np.float64) -> DeviceMatrix:
    """The reference's GENMAT (genmat.c:106-128): X(i, j) the (j m + i + 1)-th drand48 draw."""
    X = DeviceMatrix(m, n, dtype)
    _check("linompss_gen_drand48", lib.linompss_gen_drand48(
        0 if np.dtype(dtype) == np.float64 else 1, m, n, seed, X.ptr, X.ld), False)
    return X


def relerr_lower_dev(dL: DeviceMatrix, dA: DeviceMatrix) -> float:
    """The dchol check (chol_check.c:12-42): DMAT_RELERR of L L^T against A, lower triangles."""
    out = _d(0.0)
    _check("linompss_relerr_lower_dev", lib.linompss_relerr_lower_dev(
        0 if dL.dtype == np.float64 else 1, dL.rows, dL.ptr, dL.ld, dA.ptr, dA.ld, ctypes.byref(out)), False)
    return out.value


def gen_gauss(m: int, n: int, seed: int) -> DeviceMatrix:
    G = DeviceMatrix(m, n, np.float64)
    _check("linompss_gen_gauss", lib.linompss_gen_gauss(m, n, seed, G.ptr, G.ld), False)
    return G


_sig("linompss_dberr_dev", _i, _i, _i, _p, _i64, _p, _i64, _p, _i64, ctypes.POINTER(_d))


def dberr_dev(A: DeviceMatrix, X: DeviceMatrix, B: DeviceMatrix) -> float:
    """FP64 normwise backward error of X for A X = B, computed on the device."""
    e = _d(0.0)
    _check("linompss_dberr_dev", lib.linompss_dberr_dev(A.cols, X.cols, A.ptr, A.ld, X.ptr, X.ld, B.ptr, B.ld,
                                                         ctypes.byref(e)), False)
    return e.value


def copy_dev(dst: DeviceMatrix, src: DeviceMatrix) -> None:
    if dst.nbytes < src.nbytes:
        raise ValueError("destination too small")
    _check("linompss_memcpy", lib.linompss_memcpy(dst.ptr, src.ptr, src.nbytes, 3), False)


def last_exec_stats() -> ExecStats:
    st = ExecStats()
    lib.linompss_last_exec_stats(ctypes.byref(st))
    return st


# ---- plan introspection -----------------------------------------------------

def plan_tdg(what: int, n: int, b: int, nrhs: int = 0):
    """Task graph of a plan: (types[T], coords[T,3], edges[E,2])."""
    nt, ne = _i(0), _i(0)
    lib.linompss_plan_tdg(what, n, b, nrhs, None, None, 0, None, 0, ctypes.byref(nt), ctypes.byref(ne))
    types = np.zeros(nt.value, dtype=np.int32)
    coords = np.zeros((nt.value, 3), dtype=np.int32)
    edges = np.zeros((ne.value, 2), dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(_ip)
    rc = lib.linompss_plan_tdg(what, n, b, nrhs, ip(types), ip(coords), nt.value, ip(edges), ne.value,
                               ctypes.byref(nt), ctypes.byref(ne))
    if rc != 0:
        raise LinompssError("linompss_plan_tdg", rc)
    return types, coords, edges


def plan_info(what: int, n: int, b: int, nrhs: int = 0, opts: ExecOpts | None = None):
    u, t, s = _i(0), _i(0), _d(0.0)
    rc = lib.linompss_plan_info(what, n, b, nrhs, ctypes.byref(opts) if opts is not None else None,
                                ctypes.byref(u), ctypes.byref(t), ctypes.byref(s))
    if rc != 0:
        raise LinompssError("linompss_plan_info", rc)
    return u.value, t.value, s.value


# ---- factor autotuner -------------------------------------------------------------

class FactorConfig(ctypes.Structure):
    """Tile size and work-unit shapes of a factor plan (linompss_factor_config)."""
    _fields_ = [("b", _i), ("potrf_sub", _i), ("gemm_wide_m", _i), ("gemm_wide_n", _i), ("trsm_rows_wide", _i),
                ("wide_tail", _i), ("update_group", _i), ("ms", _d)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_sig("linompss_tune_factor", _i, _i, _i, _i, _i, _i, _i, ctypes.POINTER(FactorConfig), ctypes.POINTER(FactorConfig),
     _i, _ip)
_sig("linompss_set_factor_config", _i, _i, _i, _i, _i, ctypes.POINTER(FactorConfig))
_sig("linompss_get_factor_config", _i, _i, _i, _i, _i, _i, ctypes.POINTER(FactorConfig))
_sig("linompss_clear_factor_configs", _i)

_DT = {np.dtype(np.float64): 0, np.dtype(np.float32): 1}


def tune_factor(n: int, b: int = 0, dtype=np.float64, variant: str = "rl", update_prec: int = 0, reps: int = 3):
    """Autotunes the factor of shape (dtype, variant, n, b, update_prec) on the
    device (b = 0 also searches the tile size) and records the winner, which
    every later factor of that shape uses.  Returns (best, tried): the winning
    FactorConfig and the list of timed candidates (dicts, in order)."""
    best = FactorConfig()
    tried = (FactorConfig * 64)()
    nt = _i(0)
    rc = lib.linompss_tune_factor(_DT[np.dtype(dtype)], {"rl": 0, "ll": 1}[variant], n, b, update_prec, reps,
                                  ctypes.byref(best), tried, 64, ctypes.byref(nt))
    _check("linompss_tune_factor", rc, False)
    return best, [tried[q].as_dict() for q in range(min(nt.value, 64))]


def set_factor_config(n: int, cfg: FactorConfig, dtype=np.float64, variant: str = "rl", update_prec: int = 0):
    _check("linompss_set_factor_config",
           lib.linompss_set_factor_config(_DT[np.dtype(dtype)], {"rl": 0, "ll": 1}[variant], n, update_prec,
                                          ctypes.byref(cfg)), False)


def get_factor_config(n: int, b: int, dtype=np.float64, variant: str = "rl", update_prec: int = 0):
    """(recorded, FactorConfig): the configuration a factor of this shape uses."""
    cfg = FactorConfig()
    rc = lib.linompss_get_factor_config(_DT[np.dtype(dtype)], {"rl": 0, "ll": 1}[variant], n, b, update_prec,
                                        ctypes.byref(cfg))
    if rc < 0:
        raise LinompssError("linompss_get_factor_config", rc)
    return rc == 1, cfg


def clear_factor_configs() -> None:
    lib.linompss_clear_factor_configs()


# ---- batched systems -----------------------------------------------------------

BATCH_NMAX = 512


class BatchOpts(ctypes.Structure):
    _fields_ = [("max_sweeps", _i), ("tol", _d), ("stall_ratio", _d), ("wg_per_cu", _i), ("order", _i),
                ("use_graph", _i), ("panel", _i)]


class BatchStatus(ctypes.Structure):
    _fields_ = [("sweeps", _i), ("fallback", _i), ("info", _i), ("reserved", _i), ("berr", _d)]


BATCH_STATUS_DTYPE = np.dtype([("sweeps", "i4"), ("fallback", "i4"), ("info", "i4"), ("reserved", "i4"),
                               ("berr", "f8")])

_pp = ctypes.POINTER(_p)
_sig("linompss_batch_plan_create", _i, _i, _ip, _pp, ctypes.POINTER(_i64), _pp, _pp, ctypes.POINTER(BatchOpts),
     _pp)
_sig("linompss_batch_run", _i, _p, _p, _p)
_sig("linompss_batch_plan_info", _i, _p, _ip, _ip, _ip, ctypes.POINTER(_d))
_sig("linompss_batch_plan_destroy", _i, _p)
_sig("linompss_batch_plan_panels", _i, _p, ctypes.POINTER(ctypes.c_uint))
_sig("linompss_dposv_ir_batched_dev", _i, _i, _ip, _pp, ctypes.POINTER(_i64), _pp, _pp, _p,
     ctypes.POINTER(BatchOpts), _p)


class DeviceBatch:
    """Systems packed back to back in device memory: A_i (n_i x n_i, column-major,
    lda = n_i), b_i and x_i (n_i each)."""

    def __init__(self, sizes):
        self.sizes = np.asarray(sizes, dtype=np.int32)
        if self.sizes.ndim != 1 or (self.sizes < 0).any():
            raise ValueError("sizes must be a 1-D array of non-negative orders")
        n64 = self.sizes.astype(np.int64)
        self.offA = np.concatenate([[0], np.cumsum(n64 * n64)])
        self.offV = np.concatenate([[0], np.cumsum(n64)])
        self.A = DeviceMatrix(max(1, int(self.offA[-1])), 1)
        self.B = DeviceMatrix(max(1, int(self.offV[-1])), 1)
        self.X = DeviceMatrix(max(1, int(self.offV[-1])), 1)
        cnt = len(self.sizes)
        self._n = (_i * cnt)(*self.sizes.tolist())
        self._lda = (_i64 * cnt)(*np.maximum(n64, 1).tolist())
        self._pa = (_p * cnt)(*(self.A.ptr + 8 * self.offA[:-1]).tolist())
        self._pb = (_p * cnt)(*(self.B.ptr + 8 * self.offV[:-1]).tolist())
        self._px = (_p * cnt)(*(self.X.ptr + 8 * self.offV[:-1]).tolist())

    @property
    def count(self) -> int:
        return len(self.sizes)

    def upload(self, mats, rhs) -> None:
        """mats[i]: n_i x n_i (only the lower triangle is read); rhs[i]: n_i."""
        hA = np.empty(max(1, int(self.offA[-1])))
        hB = np.empty(max(1, int(self.offV[-1])))
        for i, n in enumerate(self.sizes):
            hA[self.offA[i]:self.offA[i + 1]] = np.asarray(mats[i], dtype=np.float64).ravel(order="F")
            hB[self.offV[i]:self.offV[i + 1]] = np.asarray(rhs[i], dtype=np.float64).ravel()
        self.A.upload(hA.reshape(-1, 1))
        self.B.upload(hB.reshape(-1, 1))

    def solutions(self) -> list:
        h = self.X.to_host().ravel()
        return [h[self.offV[i]:self.offV[i + 1]].copy() for i in range(self.count)]

    def free(self) -> None:
        for m in (self.A, self.B, self.X):
            m.free()


class BatchPlan:
    """A planned batch (linompss_batch_plan_create): schedule chosen (or
    autotuned on the batch), launch captured as a hipGraph; run() replays it."""

    def __init__(self, batch: DeviceBatch, opts: BatchOpts | None = None):
        self.batch = batch
        self._h = _p()
        rc = lib.linompss_batch_plan_create(batch.count, batch._n, batch._pa, batch._lda, batch._pb, batch._px,
                                            ctypes.byref(opts) if opts is not None else None,
                                            ctypes.byref(self._h))
        _check("linompss_batch_plan_create", rc, False)

    def run(self, stream=None) -> np.ndarray:
        st = np.zeros(self.batch.count, dtype=BATCH_STATUS_DTYPE)
        _check("linompss_batch_run", lib.linompss_batch_run(self._h, _s(stream), st.ctypes.data_as(_p)), False)
        return st

    def info(self) -> dict:
        g, w, o, t = _i(0), _i(0), _i(0), _d(0.0)
        _check("linompss_batch_plan_info", lib.linompss_batch_plan_info(self._h, ctypes.byref(g), ctypes.byref(w),
                                                                        ctypes.byref(o), ctypes.byref(t)), False)
        m = ctypes.c_uint(0)
        _check("linompss_batch_plan_panels", lib.linompss_batch_plan_panels(self._h, ctypes.byref(m)), False)
        return {"grid": g.value, "wg_per_cu": w.value, "order": o.value, "tune_ms": t.value,
                "panel_mask": m.value}

    def close(self) -> None:
        if self._h:
            lib.linompss_batch_plan_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dposv_ir_batched_dev(batch: DeviceBatch, opts: BatchOpts | None = None, stream=None) -> np.ndarray:
    """One-shot batched refinement solve (plan, run, destroy); returns the statuses."""
    st = np.zeros(batch.count, dtype=BATCH_STATUS_DTYPE)
    rc = lib.linompss_dposv_ir_batched_dev(batch.count, batch._n, batch._pa, batch._lda, batch._pb, batch._px,
                                           _s(stream), ctypes.byref(opts) if opts is not None else None,
                                           st.ctypes.data_as(_p))
    _check("linompss_dposv_ir_batched_dev", rc, False)
    return st


def gen_batch_spd(count: int, nmin: int, nmax: int, seed: int, shift: float = 0.1):
    """Host workload: orders uniform in [nmin, nmax], A_i = G_i G_i^T / n_i + shift I
    (G_i i.i.d. N(0, 1)), b_i uniform in [-1, 1); numpy PCG64 stream `seed`."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count).astype(np.int32)
    mats, rhs = [], []
    for n in sizes:
        G = rng.standard_normal((n, n))
        A = G @ G.T
        A /= max(int(n), 1)
        A[np.diag_indices(n)] += shift
        mats.append(np.asfortranarray(A))
        rhs.append(rng.uniform(-1.0, 1.0, n))
    return sizes, mats, rhs
