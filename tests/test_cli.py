"""The CLI front ends' argument handling on the host (no device work).

The programs mirror the reference's dchol_{rl,ll} / schol_{rl,ll}
(src/cholesky/dense/chol.c:40-84, chol_config.c:17-47) and dspdsolv /
sspdsolv (src/spdsolv/spdsolv.c:36-63, spdsolv_config.c:14-40): usage text
and exit 1 below the required argument count, the padding warning for m not
a multiple of b, exit 2 when the matrices cannot be set up (here: no device).
"""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parents[1] / "linompss_amd" / "bin"
CHOL = ["dchol_rl", "dchol_ll", "schol_rl", "schol_ll"]
SOLV = ["dspdsolv", "sspdsolv"]


def run(prog, *args, cwd=None):
    return subprocess.run([str(BIN / prog), *map(str, args)], capture_output=True, text=True, timeout=60, cwd=cwd)


@pytest.mark.parametrize("prog", CHOL + SOLV)
def test_built(prog):
    assert (BIN / prog).exists(), f"{prog} not built (linompss_amd/build.py)"


@pytest.mark.parametrize("prog", CHOL)
def test_chol_usage(prog):
    for args in ([], [64]):
        p = run(prog, *args)
        assert p.returncode == 1
        assert p.stdout == f"usage: {BIN / prog} n ts [bs] [reps] [check]\n"


@pytest.mark.parametrize("prog", SOLV)
def test_spdsolv_usage(prog):
    for args in ([], [64], [64, 1]):
        p = run(prog, *args)
        assert p.returncode == 1
        assert p.stdout == f"usage: {BIN / prog} m n b [reps] [check]\n"


def test_setup_failure_without_device(tmp_path):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a device is present: setup succeeds")
    p = run("dchol_rl", 64, 16, 16, 1, 1, cwd=tmp_path)
    assert p.returncode == 2 and p.stderr == "err: allocating matrix\n"
    p = run("dspdsolv", 1000, 2, 256, cwd=tmp_path)
    assert p.returncode == 2
    assert p.stdout == "warn: padding m to 1024\n"
    assert not (tmp_path / "ompss-chol.stats").exists()
