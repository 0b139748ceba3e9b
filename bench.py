#!/usr/bin/env python3
"""SPD-solve benchmark: GFLOP/s at n = 32768 to FP64 accuracy, mixed-precision
iterative refinement vs the plain FP64 factor + solve, on one MI355X.

Workload (the refinement configuration of the tiled Cholesky path):
  n = 32768, tile b = 512, A = G G^T / n + alpha I with G i.i.d. N(0,1)
  (Philox, seed 1) and alpha set so that cond2(A) ~ 1e4, right-hand sides
  uniform in [-1, 1).  One step = one solve to a normwise backward error
  <= 1e-14: FP32 tiled factor (one persistent launch), FP32 sweeps, FP64
  residual / update / test, FP64-factor fallback if refinement stalls.
  GFLOP/s = (n^3 / 3) / time-to-solution, inputs resident in HBM.

Beside it: the plain FP64 factor + sweeps (one persistent launch), the
roofline of the dominant kernel (the FP32 factor launch, FP32 matrix peak),
and the CPU baseline: the oracle's OpenMP-task restatement of the reference
loop nest (sequential BLAS per tile task) on a bounded sample.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SPD solve GFLOP/s at n=32768 to FP64 accuracy: mixed-prec IR vs FP64 POTRF"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "hbm": 8000.0}  # TFLOP/s, GB/s (MI355X)


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n, b, threads):
    """The oracle's OpenMP-task FP64 RL factor + sweeps on an n x n sample."""
    os.environ.setdefault("OMP_MAX_TASK_PRIORITY", "10000")
    import numpy as np
    import oracle
    A = oracle.genmat_spd(n, 0, diag_mode=1)
    B = np.asfortranarray(A @ np.ones((n, 1)))
    t0 = oracle.wtime()
    info = oracle.dspdsolv(A, B, b, threads)
    t = oracle.wtime() - t0
    ok = info == 0 and float(abs(B - 1.0).max()) < 1e-8
    return {"value": round(n ** 3 / 3.0 / t / 1e9, 2), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"FP64 OpenMP-task RL factor + forward/backward sweeps (oracle restatement of "
                      f"chol_rlmain.c + bblas_dltrsm over sequential MKL tile calls), n={n}, b={b}, "
                      f"1 RHS, {t:.2f} s, correct={ok}"}


def traffic_from(fetch_csv, write_csv, out, n):
    """HBM traffic of the FP32 factor launch from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what the
    same counter reads for the device copy of A (lx_copy_kernel: n*n*8 bytes
    read and written) in the same run."""
    import csv
    known_kib = n * n * 8 / 1024.0

    def read(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copy = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"].startswith("lx_copy_kernel")]
        fac = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"].startswith("void lx_exec_kernel<float, 0>")]
        if not copy or not fac:
            raise SystemExit(f"{path}: no lx_copy_kernel / FP32 factor dispatch with {counter}")
        return max(copy), sum(fac) / len(fac), len(fac)

    fc, ff, nf = read(fetch_csv, "FETCH_SIZE")
    wc, wf, nw = read(write_csv, "WRITE_SIZE")
    fs, ws = known_kib / fc, known_kib / wc
    res = {
        "kernel": "lx_exec_kernel<float, 0> (FP32 tiled factor, one persistent launch)",
        "workload": f"n={n}, b=512, FP32 factor inside the mixed-precision solve",
        "copy_calibration": {"bytes": n * n * 8, "fetch_kib": fc, "write_kib": wc,
                             "fetch_scale": fs, "write_scale": ws},
        "factor_fetch_kib": ff, "factor_write_kib": wf, "dispatches": [nf, nw],
        "hbm_read_bytes": ff * fs * 1024, "hbm_write_bytes": wf * ws * 1024,
        "hbm_bytes_per_launch": round((ff * fs + wf * ws) * 1024),
        "commands": [
            "rocprofv3 --pmc FETCH_SIZE -d <dir> -o run --output-format csv -- python3 bench.py --steps 1 "
            "--warmup 0 --fp64-steps 1 --no-cpu",
            "rocprofv3 --pmc WRITE_SIZE -d <dir> -o run --output-format csv -- python3 bench.py --steps 1 "
            "--warmup 0 --fp64-steps 1 --no-cpu",
            "python3 bench.py --traffic-from <fetch>/run_counter_collection.csv "
            "<write>/run_counter_collection.csv --traffic-out profiles/traffic_fp32_factor_n32768.json"],
    }
    Path(out).write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--b", type=int, default=512)
    ap.add_argument("--kappa", type=float, default=1e4)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--update-prec", type=int, default=0, choices=(0, 1),
                    help="FP32 factor's trailing SYRK/GEMM: 0 FP32 matrix cores, 1 split-BF16 matrix cores")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="derive the factor launch's HBM traffic from two rocprofv3 --pmc CSVs and exit")
    ap.add_argument("--traffic-out", default=str(ROOT / "profiles" / "traffic_fp32_factor_n32768.json"))
    args = ap.parse_args()
    if args.traffic_from:
        traffic_from(args.traffic_from[0], args.traffic_from[1], args.traffic_out, args.n)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the tiled Cholesky path here runs on a single MI355X "
              "(multi-GPU sharding is a design note only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import linompss_amd as lx

    if lx.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(3)
    n, b, nr = args.n, args.b, args.nrhs
    flops = n ** 3 / 3.0

    log(f"generating A (n={n}, cond2~{args.kappa:g}, seed {args.seed}) and B (nrhs={nr})")
    t0 = time.perf_counter()
    A, alpha, lmax = lx.gen_gauss_spd(n, args.seed, args.kappa)
    B = lx.gen_uniform(n, nr, args.seed + 100)
    X = lx.DeviceMatrix(n, nr)
    lx.lib.linompss_dev_sync()
    log(f"generated in {time.perf_counter() - t0:.1f} s: lambda_max={lmax:.6f} alpha={alpha:.3e} "
        f"cond2~{(lmax + alpha) / alpha:.3g}")

    # ---- mixed-precision refinement (the metric) ----
    ir_opts = lx.IrOpts(30, 1e-14, 0.5, args.update_prec)
    for w in range(args.warmup):
        st = lx.dposv_ir_dev(A, B, X, b, ir_opts)
        log(f"warmup {w}: sweeps={st.sweeps} fallback={st.fallback} berr={st.berr:.2e} "
            f"factor={st.t_factor_s * 1e3:.1f} ms total={st.t_total_s * 1e3:.1f} ms")
    lx.lib.linompss_dev_sync()
    t_fac, t_launch, sweeps, berrs, fallbacks = [], [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = lx.dposv_ir_dev(A, B, X, b, ir_opts)
        t_fac.append(st.t_factor_s)
        # HIP events around the FP32 factor launch, on its stream (the last
        # executor launch of the solve; the sweeps and residuals use other kernels)
        t_launch.append(lx.last_exec_stats().exec_ms * 1e-3)
        sweeps.append(st.sweeps)
        berrs.append(st.berr)
        fallbacks.append(st.fallback)
    lx.lib.linompss_dev_sync()
    t_ir = (time.perf_counter() - t0) / args.steps
    berr_ir = lx.dberr_dev(A, X, B)
    log(f"mixed: {t_ir * 1e3:.1f} ms/solve, factor {np.mean(t_fac) * 1e3:.1f} ms, sweeps {sweeps}, "
        f"berr {berr_ir:.2e}, fallback {fallbacks}")

    # ---- plain FP64 factor + solve (one persistent launch) ----
    A2 = lx.DeviceMatrix(n, n)
    X2 = lx.DeviceMatrix(n, nr)
    t64, berr64 = [], None
    for k in range(max(1, args.fp64_steps)):
        lx.copy_dev(A2, A)
        lx.copy_dev(X2, B)
        lx.lib.linompss_dev_sync()
        t0 = time.perf_counter()
        info = lx.posv_dev(A2, X2, b)
        lx.lib.linompss_dev_sync()
        t64.append(time.perf_counter() - t0)
        st64 = lx.last_exec_stats()
        log(f"fp64 posv {k}: {t64[-1] * 1e3:.1f} ms (launch {st64.exec_ms:.1f} ms, units {st64.units}) info={info}")
    berr64 = lx.dberr_dev(A, X2, B)
    t_fp64 = min(t64)
    A2.free()
    X2.free()

    t_kernel = float(np.mean(t_launch))
    roofline = {
        "bound": "mfma",
        "achieved": round(flops / t_kernel / 1e12, 3),
        "peak": PEAK["fp32"],
        "unit": "TFLOP/s",
        "frac": round(flops / t_kernel / 1e12 / PEAK["fp32"], 4),
        "traffic": None,
    }
    tf = ROOT / "profiles" / "traffic_fp32_factor_n32768.json"
    if tf.exists():
        try:
            roofline["traffic"] = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            pass

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline: n={args.cpu_n} on {threads} threads")
        try:
            cpu = cpu_baseline(args.cpu_n, b, threads)
            log(f"CPU baseline {cpu['value']} GFLOP/s")
        except Exception as e:  # the oracle needs the host BLAS it was built against
            log(f"CPU baseline unavailable: {e}")

    out = {
        "metric": METRIC,
        "value": round(flops / t_ir / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_ir * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 factor (trailing SYRK/GEMM on bf16 matrix cores, split operands hi+mid+lo, f32 "
                  "accumulation) + f64 refinement") if args.update_prec == 1 else "f32 factor + f64 refinement",
        "data": "synthetic: A = G G^T/n + alpha I, G ~ N(0,1) (Philox4x32-10, seed %d), B ~ U[-1,1)" % args.seed,
        "config": {"workload": "FP32-factor + FP64 iterative refinement, SPD solve to backward error 1e-14",
                   "update_prec": ["fp32", "split-bf16"][args.update_prec],
                   "n": n, "b": b, "cond2": args.kappa, "nrhs": nr, "alpha": alpha, "lambda_max": lmax},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "berr": berr_ir,
        "sweeps": int(np.max(sweeps)),
        "fallback": int(np.max(fallbacks)),
        "factor_ms": round(float(np.mean(t_fac)) * 1e3, 3),
        "factor_launch_ms": round(t_kernel * 1e3, 3),
        "fp64_posv_ms": round(t_fp64 * 1e3, 3),
        "fp64_posv_gflops": round(flops / t_fp64 / 1e9, 2),
        "fp64_berr": berr64,
        "speedup_vs_fp64": round(t_fp64 / t_ir, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
