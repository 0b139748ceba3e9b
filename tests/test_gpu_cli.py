"""The CLI front ends on the device, and the device generators and check they use.

* ``linompss_gen_drand48_spd`` / ``linompss_gen_drand48`` equal the oracle's
  GENMAT_SPD / GENMAT (genmat.c:49-70, 106-128, drand48 restated from POSIX)
  bit for bit, FP64 and FP32.
* ``linompss_relerr_lower_dev`` (chol_check.c:12-42 with DMAT_RELERR,
  densutil.c:50-99) against the same formula in numpy, on a factor whose
  matrix carries a known perturbation (so the value is far above rounding).
* ``dchol_rl`` / ``dchol_ll`` / ``schol_rl`` and ``dspdsolv`` / ``sspdsolv``
  on the plain FP64 workload's size (n = 4096, b = 256, seed 0): the check
  passes and exits 0, the check value printed equals the library's own for
  the same matrix, ``ompss-<app>.stats`` has the reference's four lines
  (file_log.c:43-51), a failed check exits 1.
"""
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "linompss_amd" / "bin"


def run(prog, *args, cwd, seed=0, tol=None):
    env = dict(os.environ, LINOMPSS_SEED=str(seed))
    if tol is not None:
        env["LINOMPSS_CHECK_TOL"] = str(tol)
    print(f"[cli] {prog} {' '.join(map(str, args))} seed={seed} tol={tol}", flush=True)
    return subprocess.run([str(BIN / prog), *map(str, args)], capture_output=True, text=True, timeout=300,
                          cwd=cwd, env=env)


def stats(path, reps):
    lines = path.read_text().splitlines()
    assert len(lines) == 4
    u = os.uname()
    assert lines[0] == f"{u.nodename} {u.sysname} {u.release} {u.machine}"
    m = re.fullmatch(r"time: (\d+\.\d{5}) ms", lines[1])
    assert m and float(m.group(1)) > 0
    assert lines[2] == f"reps: {reps}"
    t = re.fullmatch(r"threads: (\d+)", lines[3])
    assert t and int(t.group(1)) > 0
    return float(m.group(1))


@pytest.mark.parametrize("n,seed,pad", [(1, 0, 0), (300, 5, 3), (1024, 1733300000, 0)])
def test_gen_drand48_spd_bit_exact(lx, orc, dev, n, seed, pad):
    A = lx.gen_drand48_spd(n, seed, ld=n + pad).to_host()
    assert np.array_equal(A, orc.genmat_spd(n, seed, diag_mode=0))
    # FP32 (sgenmat_spd): each draw rounded to float, n added in float on the diagonal
    d = orc.drand48_fill(seed, n * (n + 1) // 2).astype(np.float32)
    ref = np.zeros((n, n), np.float32)
    k = 0
    for j in range(n):
        ref[j:, j] = d[k:k + n - j]
        k += n - j
    ref = np.tril(ref) + np.tril(ref, -1).T
    ref[np.diag_indices(n)] = ref[np.diag_indices(n)] + np.float32(n)
    assert np.array_equal(lx.gen_drand48_spd(n, seed, dtype=np.float32).to_host(), ref)


@pytest.mark.parametrize("m,n,seed", [(1, 1, 0), (257, 3, 7), (4096, 2, 0)])
def test_gen_drand48_bit_exact(lx, orc, dev, m, n, seed):
    assert np.array_equal(lx.gen_drand48(m, n, seed).to_host(), orc.genmat(m, n, seed))
    assert np.array_equal(lx.gen_drand48(m, n, seed, dtype=np.float32).to_host(),
                          orc.genmat(m, n, seed).astype(np.float32))


def _relerr_np(L, A):
    C = np.tril(L) @ np.tril(L).T
    a = np.abs(np.tril(C))
    d = np.abs(a - np.abs(np.tril(A)))
    return d.sum(axis=0).max() / a.sum(axis=0).max()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_relerr_lower_dev(lx, orc, dev, dt):
    n = 700
    print(f"[relerr] n={n} {np.dtype(dt).name}", flush=True)
    S = orc.genmat_spd(n, 3, diag_mode=0)
    L = np.linalg.cholesky(S)
    rng = np.random.default_rng(1)
    E = np.tril(rng.uniform(-1, 1, (n, n))) * 1e-3
    Lp = np.asfortranarray(L + np.triu(rng.uniform(-1, 1, (n, n)), 1))  # the upper triangle is ignored
    A = np.asfortranarray(S + E)
    got = lx.relerr_lower_dev(lx.DeviceMatrix.from_host(Lp.astype(dt)), lx.DeviceMatrix.from_host(A.astype(dt)))
    ref = _relerr_np(L.astype(dt).astype(np.float64), A.astype(dt).astype(np.float64))
    assert ref > 1e-6
    assert abs(got - ref) <= (1e-9 if dt == np.float64 else 1e-3) * ref


@pytest.mark.parametrize("prog,variant", [("dchol_rl", "rl"), ("dchol_ll", "ll")])
def test_dchol_cli(lx, dev, tmp_path, prog, variant):
    n, b, reps = 4096, 256, 3
    p = run(prog, n, b, b, reps, 1, cwd=tmp_path)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stderr
    m = re.fullmatch(r"check: potrf\.\.\.norm of difference: +(\S+)\n", p.stdout)
    assert m, p.stdout
    val = float(m.group(1))
    assert val <= 1e-14
    stats(tmp_path / "ompss-chol.stats", reps)
    # the same factor through the library, same matrix: the same check value, to the digit
    dA = lx.gen_drand48_spd(n, 0)
    dA0 = lx.gen_drand48_spd(n, 0)
    lx.chol_dev(dA, b, variant=variant)
    assert m.group(1) == "%.16g" % lx.relerr_lower_dev(dA, dA0)


def test_schol_cli(dev, tmp_path):
    p = run("schol_rl", 4096, 256, 256, 2, 1, cwd=tmp_path)
    assert p.returncode == 0, p.stderr
    val = float(p.stdout.split(":")[-1])
    assert 0 < val <= 1e-6
    stats(tmp_path / "ompss-chol.stats", 2)


def test_chol_cli_failed_check_exits_1(dev, tmp_path):
    """A check value above the threshold exits 1 (the threshold is tightened
    from the host; the factor itself is correct) and still writes the stats."""
    p = run("dchol_rl", 512, 128, 128, 1, 1, cwd=tmp_path, tol=1e-300)
    assert p.returncode == 1
    assert "norm of difference" in p.stdout
    stats(tmp_path / "ompss-chol.stats", 1)
    p = run("dchol_rl", 512, 128, 128, 1, 0, cwd=tmp_path, tol=1e-300)  # no check: exit 0
    assert p.returncode == 0 and p.stdout == ""


@pytest.mark.parametrize("prog,tol", [("dspdsolv", 1e-12), ("sspdsolv", 1e-4)])
def test_spdsolv_cli(dev, tmp_path, prog, tol):
    p = run(prog, 4096, 3, 256, 2, 1, cwd=tmp_path)
    print(p.stdout, p.stderr, flush=True)
    m = re.fullmatch(r"check: residual norm = (\S+)\n", p.stdout)
    assert m, p.stdout
    val = float(m.group(1))
    assert val <= tol
    assert p.returncode == (0 if val <= 1e-6 else 1)
    stats(tmp_path / "ompss-spdsolv.stats", 2)


def test_spdsolv_cli_padding_and_failure(dev, tmp_path):
    p = run("dspdsolv", 1000, 2, 256, 1, 1, cwd=tmp_path)
    assert p.returncode == 0, p.stderr
    assert p.stdout.startswith("warn: padding m to 1024\ncheck: residual norm = ")
    p = run("dspdsolv", 512, 1, 128, 1, 1, cwd=tmp_path, tol=1e-300)
    assert p.returncode == 1
    stats(tmp_path / "ompss-spdsolv.stats", 1)
