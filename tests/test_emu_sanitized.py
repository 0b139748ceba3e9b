"""The emulated kernels (tests/cpu_emu) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on a representative set of the emulator's
cases: the executor (FP64, split-BF16 and FP16-plane update units, the fused
single launch, the hypermatrix form, the bounded-wait abort), the sweep
kernels (1 to 16 right-hand sides, the mailbox's NaN case, the abort), the
batched kernel with its fallback, and the adjoint's diagonal and GEMM
kernels.  Host code only: the sanitizer runtime is preloaded into a child
Python process that loads the sanitized build.  A positive control (a heap
read one past the end) shows the sanitizer is live."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EMU = ROOT / "tests" / "cpu_emu"
CXX = "/opt/rocm/llvm/bin/clang++"

CASES = [
    "test_factor_fp64[37-16-0]", "test_factor_fp64[8-4-1]",
    "test_gemm_unit_split_bf16[6-77-200-33-0]", "test_gemm_unit_split_bf16[14-77-200-33-0]",
    "test_gemm_unit_from_planes[11-96-200-16-40]", "test_posv_single_launch[37-16-2]",
    "test_factor_hypermatrix[1-0-2-80]",
    "test_sweep_kernels[100-3-float64]", "test_sweep_kernels[300-7-float32]", "test_sweep_kernels[520-16-float64]",
    "test_batch_kernel_fallback", "test_adjoint_diag_kernel[37]", "test_adjoint_gemm_kernel[1-0]",
    "test_gemm64_float_and_upper[1-0-2]", "test_executor_bounded_wait_aborts[float64]",
    "test_sweep_bounded_wait_aborts[1]", "test_sweep_nan_with_mailbox_bits[0-float32-uint32]",
]


def _env():
    rt = subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True,
                        check=True).stdout.strip()
    if not Path(rt).exists():
        pytest.skip("no AddressSanitizer runtime")
    # -O0: the sanitized build takes seconds instead of minutes
    flags = "-O0 -g -std=c++17 -fPIC -pthread -I. -I../../linompss_amd/csrc -I../../include -Wno-unused-result -Wno-psabi"
    subprocess.run(["make", "-s", "-C", str(EMU), "asan", f"CXXFLAGS={flags}"], check=True)
    env = dict(os.environ, LX_EMU_SANITIZE="1", LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    return env


def test_sanitizer_is_live():
    env = _env()
    code = ("import ctypes; E = ctypes.CDLL('%s'); print(E.lx_emu_heap_probe(3), flush=True); E.lx_emu_heap_probe(8)"
            % (EMU / "_build" / "liblx_emu_asan.so"))
    p = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert p.stdout.strip() == "0"
    assert p.returncode != 0 and "heap-buffer-overflow" in p.stderr


def test_emulated_kernels_clean_under_sanitizers():
    env = _env()
    ids = [f"tests/test_emu.py::{c}" for c in CASES]
    p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *ids], env=env,
                       capture_output=True, text=True, cwd=str(ROOT), timeout=1200)
    out = p.stdout + p.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert p.returncode == 0, out[-4000:]
    assert f"{len(CASES)} passed" in out
