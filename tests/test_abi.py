"""The C ABI: every symbol declared in include/linompss.h is exported by the
library, argument errors are reported LAPACK-style without touching a device,
and the planner works on the host."""
import ctypes
import re
from pathlib import Path

import numpy as np

HDR = Path(__file__).resolve().parents[1] / "include" / "linompss.h"


def declared_functions():
    txt = re.sub(r"/\*.*?\*/", "", HDR.read_text(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+([a-z_][a-z0-9_]*)\s*\(", txt, re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_symbols_exported(lx):
    names = declared_functions()
    assert "ompss_dchol_rl" in names and "dspdsolv_main" in names and "linompss_dposv_ir_dev" in names
    for n in names:
        assert hasattr(lx.lib, n), n


def test_reference_entry_points_present(lx):
    # the reference's exported column-major API (ompss_dense_chol.h:32-38) and the spdsolv driver
    for n in ("ompss_dchol_rl", "ompss_dchol_ll", "ompss_schol_rl", "ompss_schol_ll", "dspdsolv_main",
              "sspdsolv_main"):
        assert hasattr(lx.lib, n)


def test_argument_errors_lapack_style(lx):
    A = np.zeros((4, 4), order="F")
    p = A.ctypes.data_as(ctypes.c_void_p)
    assert lx.lib.ompss_dchol_rl(-1, 2, 0, p, 4) == -1
    assert lx.lib.ompss_dchol_rl(4, 0, 0, p, 4) == -2
    assert lx.lib.ompss_dchol_ll(4, 2, 0, p, 3) == -5
    assert lx.lib.ompss_dchol_rl(0, 2, 0, p, 1) == 0        # empty matrix: nothing to do
    assert lx.lib.dspdsolv_main(4, 1, 0, p, p) == -3
    assert lx.lib.linompss_dltrsm(b"U", b"N", b"N", 2, 1, 4, 1, 1.0, p, 4, p, 4) == -1
    assert lx.lib.linompss_dltrsm(b"L", b"N", b"U", 2, 1, 4, 1, 1.0, p, 4, p, 4) == -3


def test_version_and_plan(lx):
    assert b"gfx950" in lx.lib.linompss_version()
    u, t, s = lx.plan_info(0, 4096, 256)
    assert t == 816 and u >= t and s > 0
    u2, t2, _ = lx.plan_info(2, 4096, 256, 1)
    assert t2 == 816 + 2 * (16 + 120)


def test_package_does_not_import_torch():
    import subprocess
    import sys
    code = "import linompss_amd, sys; print('torch' in sys.modules)"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         cwd=str(HDR.parents[1]), check=True).stdout.strip()
    assert out == "False"


def test_reference_descriptors_exported(lx):
    """OMPSSBLAS_* / OMPSSLAPACK_* (src/fp/fpmatr.c:36-47): the Fortran-BLAS
    flavour, one-letter strings; ROWMAJOR / COLMAJOR are zero there too."""
    want = {"TRANSP": b"T", "NTRANSP": b"N", "UPPERTRIANG": b"U", "LOWERTRIANG": b"L", "NTRIANG": b"x",
            "DIAGUNIT": b"U", "NDIAGUNIT": b"N", "LEFT": b"L", "RIGHT": b"R"}
    for k, v in want.items():
        assert ctypes.string_at(lx.exported_desc("OMPSSBLAS_" + k)) == v
    assert lx.exported_desc("OMPSSBLAS_ROWMAJOR") == 0 and lx.exported_desc("OMPSSBLAS_COLMAJOR") == 0
    assert ctypes.string_at(lx.exported_desc("OMPSSLAPACK_UPPERTRIANG")) == b"U"
    assert ctypes.string_at(lx.exported_desc("OMPSSLAPACK_LOWERTRIANG")) == b"L"


def test_reference_blas_flag_decoding(lx, capfd):
    """Both descriptor flavours decode to the same flag, and the argument
    checks that follow run before any device work: a bad flag or size is the
    reference's BLAS argument number (ompss_dgemm returns it; the void
    functions print it)."""
    A = np.zeros((4, 4), order="F")
    p = A.ctypes.data_as(ctypes.c_void_p)
    for fl in ("enum", "str"):
        N, T = lx.blas_desc("trans", "N", fl), lx.blas_desc("trans", "T", fl)
        assert lx.lib.ompss_dgemm(N, T, 1, 1, 1, -1, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4) == -6
        assert lx.lib.ompss_dgemm(T, N, 1, 1, 1, 4, 4, 5, 1.0, p, 4, p, 5, 0.0, p, 4) == -11
        assert lx.lib.ompss_sgemm(N, N, 1, 1, 1, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 3) == -16
    assert lx.lib.ompss_dgemm(0, 111, 1, 1, 1, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4) == -1      # no flag at all
    assert lx.lib.ompss_dgemm(111, 121, 1, 1, 1, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4) == -2    # uplo as trans
    assert lx.lib.ompss_dgemm(lx.blas_desc("uplo", "U", "str"), 111, 1, 1, 1, 4, 4, 4, 1.0, p, 4, p, 4, 0.0,
                              p, 4) == -1
    capfd.readouterr()
    lx.lib.ompss_dsyrk(lx.blas_desc("trans", "N"), 111, 1, 4, 4, 1.0, p, 4, 0.0, p, 4)
    lx.lib.ompss_dtrsm(141, 122, 111, 131, 4, 4, 0, 1.0, p, 4, p, 4)
    lx.lib.ompss_strsm(lx.exported_desc("OMPSSBLAS_RIGHT"), lx.exported_desc("OMPSSBLAS_LOWERTRIANG"),
                       lx.exported_desc("OMPSSBLAS_TRANSP"), lx.exported_desc("OMPSSBLAS_NTRIANG"), 4, 4, 1, 1.0,
                       p, 4, p, 4)
    lx.lib.task_dpotrf(ord("X"), 4, p, 4, 0)
    lx.lib.task_dpotrf(lx.lapack_desc("L", "str"), 4, p, 3, 0)
    lx.lib.task_dpotrf(lx.exported_desc("OMPSSLAPACK_UPPERTRIANG"), 0, p, 1, 0)  # empty: nothing printed
    lx.lib.task_dgemm(111, 111, -2, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, 0)
    err = capfd.readouterr().err.splitlines()
    assert err == ["linompss: ompss_dsyrk returned -1", "linompss: ompss_dtrsm returned -7",
                   "linompss: ompss_strsm returned -4", "linompss: task_dpotrf returned -1",
                   "linompss: task_dpotrf returned -4", "linompss: task_dgemm returned -6"]
