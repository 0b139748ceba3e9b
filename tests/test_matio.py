"""Dense-matrix files in the Octave text format of the reference's
refinement app (A.dat / RHS.dat: src/itref/itref_setup.c:98-121, 155-199).

Pins: the exact text of print_dense2mm (src/matutil/matfprint.c:104-124),
restated by Python's own "%.16e" formatting; an independent parse with
numpy.loadtxt; exact round trips of FP64 (17 significant digits) and FP32;
read_mm2dense's file-order reading with '#' lines skipped and its
repeat-the-last-value behaviour on a line that does not parse
(src/matutil/matfread.c:13-27); argument and I/O errors.  Host code only, no
GPU involved.
"""
import os
import subprocess
import sys

import numpy as np
import pytest


def _text(name, A):
    m, n = A.shape
    lines = [f"# name: {name}", "# type: matrix", f"# rows: {m}", f"# columns: {n}"]
    lines += ["%.16e " % float(A[i, j]) for i in range(m) for j in range(n)]
    return "\n".join(lines) + "\n"


def test_writer_text_and_independent_parse(lx, tmp_path):
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.standard_normal((5, 3)) * 10.0 ** rng.integers(-20, 20, (5, 3)))
    f = tmp_path / "A.dat"
    lx.write_dense(f, A, "A")
    assert f.read_text() == _text("A", A)
    flat = np.loadtxt(f, comments="#")
    assert np.array_equal(flat.reshape(5, 3), A)  # values are row-major in the file


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(1, 1), (7, 1), (4, 9), (0, 3)])
def test_round_trip(lx, tmp_path, dt, shape):
    rng = np.random.default_rng(1)
    A = np.asfortranarray(rng.standard_normal(shape).astype(dt))
    f = tmp_path / "M.dat"
    lx.write_dense(f, A, "M")
    B = lx.read_dense(f, dt)
    assert B.shape == A.shape and B.dtype == dt
    assert np.array_equal(B, A)


def test_raw_reader_file_order(lx, tmp_path):
    """read_mm2dense stores values in file order (the reference's setup reads
    them straight into its column-major arrays, which transposes a
    non-symmetric matrix), up to the count it is given."""
    A = np.asfortranarray(np.arange(6.0).reshape(2, 3))
    f = tmp_path / "A.dat"
    lx.write_dense(f, A, "A")
    assert np.array_equal(lx.read_values(f, 6), [0, 1, 2, 3, 4, 5])
    assert np.array_equal(lx.read_values(f, 4), [0, 1, 2, 3])
    assert lx.read_values(f, 100).size == 6


def test_raw_reader_skips_comments_and_repeats_on_bad_lines(lx, tmp_path):
    f = tmp_path / "odd.dat"
    f.write_text("# a comment\n1.5\n# another\nnot a number\n\n-2e3 \n")
    # the unparsable line and the blank one each store the last value again
    assert np.array_equal(lx.read_values(f, 10), [1.5, 1.5, 1.5, -2000.0])
    g = tmp_path / "lead.dat"
    g.write_text("junk\n3\n")
    assert np.array_equal(lx.read_values(g, 2), [0.0, 3.0])  # nothing read yet: 0


def test_errors(lx, tmp_path):
    import ctypes
    missing = os.fsencode(tmp_path / "nope.dat")
    assert lx.lib.linompss_read_mm2ddense(missing, 3, None) == -3  # out may be null only for count 0
    out = (ctypes.c_double * 3)()
    assert lx.lib.linompss_read_mm2ddense(missing, 3, out) == -1006
    assert lx.lib.linompss_fprint_ddense2mm(None, b"A", 1, 1, out, 1) == -1
    assert lx.lib.linompss_fprint_ddense2mm(os.fsencode(tmp_path / "x.dat"), b"A", 2, 2, out, 1) == -6
    f = tmp_path / "A.dat"
    lx.write_dense(f, np.ones((2, 2)), "A")
    assert lx.lib.linompss_read_ddense2mm(os.fsencode(f), 3, 2, out, 3) == -1006  # header says 2 x 2
    f.write_text("# rows: 2\n# columns: 2\n1\n2\n3\n")  # one value short
    buf = (ctypes.c_double * 4)()
    assert lx.lib.linompss_read_ddense2mm(os.fsencode(f), 2, 2, buf, 2) == -1006


def test_build_module_loads_without_the_library(tmp_path):
    """build() must work on a fresh checkout: loading the build module may not
    import the package (whose import needs the library being built)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path.insert(0, %r); import __graft_entry__ as g; m = g._build_module(); "
            "assert hasattr(m, 'build'); assert 'linompss_amd' not in sys.modules; print('ok')" % root)
    env = dict(os.environ, LINOMPSS_LIB=str(tmp_path / "absent.so"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
