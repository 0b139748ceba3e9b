"""Builds liblinompss.so (the HIP/gfx950 library) in-tree.

The library is compiled for gfx950 only.  Device code is built with packed
FP32 instructions disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
Objects go to ``linompss_amd/_build``; the shared library is written next to
this file.  The build always recompiles (it is also the "does it build" check).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
BUILD = PKG / "_build"
LIB = PKG / "liblinompss.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-mcode-object-version=5",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep MFMA accumulators in VGPRs: avoids a VGPR<->AGPR copy of every
    # accumulator per loop iteration in the tile GEMMs
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out.strip() and os.environ.get("LINOMPSS_BUILD_VERBOSE"):
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    incs = [f"-I{INCLUDE}", f"-I{CSRC}"]
    objs = []
    jobs = []
    for src in sorted(CSRC.glob("*.cpp")):
        obj = BUILD / (src.stem + ".o")
        jobs.append(["g++", *HOST_FLAGS, *incs, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in sorted(CSRC.glob("*.hip")):
        obj = BUILD / (src.stem + ".o")
        jobs.append([HIPCC, *DEVICE_FLAGS, *incs, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out))
    if failed:
        for cmd, out in failed:
            sys.stderr.write(f"$ {' '.join(cmd)}\n{out}\n")
        raise RuntimeError("liblinompss build failed")
    tmp = LIB.with_suffix(".so.tmp")
    # -Bsymbolic: calls between this library's own exported functions (the
    # tile tasks call ompss_dchol_rl, ...) bind inside it, never to a
    # same-named symbol another library in the process exports
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic",
          f"-Wl,--version-script={CSRC / 'liblinompss.map'}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


CLI = PKG / "cli"
BIN = PKG / "bin"
# the reference's programs, built per precision / variant as it builds them
# (chol.c with -DDOUBLE_PRECISION -DUSE_RL, ...; spdsolv.c)
CLI_PROGRAMS = {
    "dchol_rl": ("lx_chol_cli.cpp", 0, 0), "dchol_ll": ("lx_chol_cli.cpp", 0, 1),
    "schol_rl": ("lx_chol_cli.cpp", 1, 0), "schol_ll": ("lx_chol_cli.cpp", 1, 1),
    "dspdsolv": ("lx_spdsolv_cli.cpp", 0, 0), "sspdsolv": ("lx_spdsolv_cli.cpp", 1, 0),
}


def build_cli() -> Path:
    """The CLI front ends (host C++ over the C-ABI only) into linompss_amd/bin,
    linked against the in-tree liblinompss.so."""
    BIN.mkdir(exist_ok=True)
    for prog, (src, dtype, variant) in CLI_PROGRAMS.items():
        tmp = BIN / (prog + ".tmp")
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{CLI}", f"-DLX_CLI_DTYPE={dtype}",
              f"-DLX_CLI_VARIANT={variant}", str(CLI / src), "-o", str(tmp), f"-L{PKG}", "-llinompss",
              "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, BIN / prog)
    return BIN


def build_oracle() -> Path:
    """Builds the CPU oracle (test infrastructure) with its Makefile."""
    odir = PKG.parent / "oracle"
    _run(["make", "-s", "-C", str(odir), "-B"])
    # the reference's own sources, compiled in place (only where they exist)
    if Path("/root/reference/src/cholesky/dense/chol_rlmain.c").exists():
        _run(["make", "-s", "-C", str(odir), "-B", "ref"])
    return odir / "_build" / "liblx_oracle.so"


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cli())
