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

  The factor's trailing SYRK/GEMM run on BF16 matrix cores with each FP32
  operand split into three BF16 parts (hi + mid + lo carry all 24 bits) and
  six partial products accumulated in FP32 (--update-prec 1, the default);
  --update-prec 0 runs them on FP32 matrix cores.

Beside it: the plain FP64 factor + sweeps (one persistent launch), the
roofline of the dominant kernel (the FP32 factor launch; with split-BF16
updates its ceiling is the BF16 peak over the six products per FP32
multiply-add), and the CPU baseline: the oracle's OpenMP-task restatement of
the reference loop nest (sequential BLAS per tile task) on a bounded sample.

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


def host_threads(requested: int) -> int:
    """CPU-baseline threads: --cpu-threads, else the host share the environment
    gives this process (OMP_NUM_THREADS), else every core up to 16."""
    if requested:
        return requested
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else min(16, os.cpu_count() or 1)


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


def cpu_baseline_batched(sizes, mats, rhs, threads, sample=512):
    """The oracle's refinement solve (dposv_ir, one tile per system) over the
    first `sample` systems, one system per host thread at a time."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    idx = list(range(min(sample, len(sizes))))

    def one(i):
        n = int(sizes[i])
        X, st, info = oracle.dposv_ir(mats[i], rhs[i].reshape(-1, 1), max(n, 1), nthreads=1)
        return info == 0 and st.berr <= 1e-14

    t0 = oracle.wtime()
    with ThreadPoolExecutor(threads) as ex:
        ok = all(ex.map(one, idx))
    t = oracle.wtime() - t0
    return {"value": round(len(idx) / t, 1), "unit": "systems/s", "cores": threads, "kind": "port",
            "sample": f"oracle refinement solve (FP32 LAPACK factor of one tile + FP64 residual/update, "
                      f"itref.c:999-1101 contract) of the first {len(idx)} systems, one system per thread, "
                      f"{t:.2f} s, all converged={ok}"}


def bench_batched(args):
    """Batched uneven SPD solves: systems/s for one persistent launch (a
    replayed hipGraph) over 4096 systems, inputs resident in HBM."""
    import numpy as np
    import linompss_amd as lx
    if lx.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(3)
    cnt = args.batch_count
    log(f"generating {cnt} systems, orders uniform in [32, 512], seed 2")
    sizes, mats, rhs = lx.gen_batch_spd(cnt, 32, 512, seed=2)
    batch = lx.DeviceBatch(sizes)
    batch.upload(mats, rhs)
    t0 = time.perf_counter()
    plan = lx.BatchPlan(batch)  # autotuned schedule, captured graph
    pinfo = plan.info()
    log(f"plan: {pinfo} ({(time.perf_counter() - t0) * 1e3:.0f} ms incl. tuning)")
    for w in range(args.warmup):
        plan.run()
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = plan.run()
    t = (time.perf_counter() - t0) / args.steps
    X = batch.solutions()
    berr_np = max(float(np.abs(rhs[i] - mats[i] @ X[i]).max() / (np.abs(mats[i]).sum(1).max() * np.abs(X[i]).max()))
                  for i in range(0, cnt, max(1, cnt // 256)))
    n64 = sizes.astype(np.float64)
    flops = float((n64 ** 3).sum() / 3.0)
    abytes = float((12.0 * n64 * n64).sum())  # FP64 A + its FP32 copy, once
    log(f"batched: {t * 1e3:.3f} ms per launch, {cnt / t:.0f} systems/s, max berr {st['berr'].max():.2e}, "
        f"sweeps max {st['sweeps'].max()}, fallbacks {int(st['fallback'].sum())}")
    cpu = None
    if not args.no_cpu:
        threads = host_threads(args.cpu_threads)
        try:
            cpu = cpu_baseline_batched(sizes, mats, rhs, threads)
            log(f"CPU baseline {cpu['value']} systems/s")
        except Exception as e:
            log(f"CPU baseline unavailable: {e}")
    out = {
        "metric": "batched uneven SPD solves to FP64 backward error 1e-14 (systems/s)",
        "value": round(cnt / t, 1),
        "unit": "systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refinement (per system)",
        "data": "synthetic: A_i = G_i G_i^T / n_i + 0.1 I, G_i ~ N(0,1), b_i ~ U[-1,1) (numpy PCG64, seed 2)",
        "config": {"workload": "batched uneven SPD solves, one persistent launch (hipGraph replay)",
                   "count": cnt, "n_min": 32, "n_max": 512, "sum_n": int(sizes.sum()), "nrhs": 1,
                   "schedule": pinfo},
        "roofline": {"bound": "hbm", "achieved": round(abytes / t / 1e9, 1), "peak": PEAK["hbm"], "unit": "GB/s",
                     "frac": round(abytes / t / 1e9 / PEAK["hbm"], 4), "traffic": None,
                     "basis": "12 n_i^2 bytes per system (FP64 A and its FP32 copy, once)"},
        "gflops": round(flops / t / 1e9, 1),
        "cpu_baseline": cpu,
        "berr_max": float(st["berr"].max()),
        "berr_max_numpy_sample": berr_np,
        "sweeps_max": int(st["sweeps"].max()),
        "fallbacks": int(st["fallback"].sum()),
    }
    print(json.dumps(out), flush=True)
    plan.close()
    batch.free()


def cpu_baseline_gp(P, y, ell, noise, threads):
    """The oracle's FP64 closed-form gradient (K^-1 through LAPACK) on the host."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import adjoint as adj
    t0 = time.perf_counter()
    dK, dell = adj.gp_nll_grad(P, y, ell, noise)
    t = time.perf_counter() - t0
    n = len(y)
    return dell, {"value": round(t * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
                  "sample": f"oracle closed-form GP gradient dK = (K^-1 - a a^T)/2 in FP64 numpy/LAPACK "
                            f"(explicit inverse), n={n}, one evaluation, {t:.2f} s"}


def bench_gp(args):
    """GP negative log marginal likelihood at n points: forward (FP64 factor,
    log det, solve) + backward (the forward's factor reused, solve adjoint,
    blocked Cholesky adjoint) through torch.autograd, K resident on the
    device."""
    import numpy as np
    import linompss_amd as lx
    if lx.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(3)
    import linompss_amd.autograd as ag
    import torch
    n, ell0, noise = args.gp_n, 0.3, 1e-2
    log(f"GP workload: {n} points in [0,1]^3 (seed 3), l={ell0}, noise={noise}")
    rng = np.random.default_rng(3)
    P = rng.uniform(0.0, 1.0, (n, 3))
    pts = torch.from_numpy(P)
    with torch.no_grad():
        K0 = ag.gp_kernel(pts, torch.tensor(ell0, dtype=torch.float64), noise).numpy()
    L0 = np.asfortranarray(K0.copy())
    lx.ompss_dchol_rl(L0, 256)
    y = np.tril(L0) @ rng.standard_normal(n)  # y ~ N(0, K)
    yt = torch.from_numpy(y)

    # the FP64 factor's tile and unit shapes for this n, chosen on the device
    # (one factor per step, in the forward pass)
    gp_b = args.gp_b
    tuned = None
    if not args.no_tune:
        t0 = time.perf_counter()
        best, tried = lx.tune_factor(n, 0, np.float64, "rl", 0, reps=3)
        gp_b = best.b
        tuned = dict(best.as_dict(), candidates=len(tried), tune_s=round(time.perf_counter() - t0, 1))
        log(f"tuned FP64 factor at n={n}: {tuned}")

    def step(tm):
        # K built on the device from the points (GPMarginalLikelihood): only the
        # points, y and scalars cross PCIe
        ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
        loss = ag.gp_nll_device(pts, yt, ell, 1.0, noise, b=gp_b, timing=tm)
        loss.backward()
        return loss.item(), ell.grad.item()

    for w in range(args.warmup):
        step({})
    dev_s, wall_s = [], []
    for k in range(args.steps):
        tm = {}
        t0 = time.perf_counter()
        nll, g = step(tm)
        wall_s.append(time.perf_counter() - t0)
        dev_s.append(tm["forward_s"] + tm["backward_s"])
        log(f"step {k}: nll={nll:.10e} dnll/dl={g:.10e} device fwd {tm['forward_s'] * 1e3:.1f} ms "
            f"bwd {tm['backward_s'] * 1e3:.1f} ms, wall {wall_s[-1] * 1e3:.0f} ms")
    # central difference (h = 1e-6) of the same forward
    h = 1e-6
    with torch.no_grad():
        fp = ag.gp_nll_device(pts, yt, ell0 + h, 1.0, noise, b=gp_b).item()
        fm = ag.gp_nll_device(pts, yt, ell0 - h, 1.0, noise, b=gp_b).item()
    fd = (fp - fm) / (2 * h)
    t_dev = float(np.median(dev_s))
    flops = 1.0 * n ** 3  # forward factor n^3/3, adjoint 2n^3/3
    cpu, dell_ref = None, None
    if not args.no_cpu:
        threads = host_threads(args.cpu_threads)
        try:
            dell_ref, cpu = cpu_baseline_gp(P, y, ell0, noise, threads)
            log(f"CPU closed-form gradient {cpu['value']} ms, dnll/dl={dell_ref:.10e}")
        except Exception as e:
            log(f"CPU baseline unavailable: {e}")
    out = {
        "metric": f"GP negative log marginal likelihood forward+backward at n={n} (device ms)",
        "value": round(t_dev * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.median(wall_s)) * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: points ~ U[0,1]^3 (numpy PCG64 seed 3), squared-exponential K, l=0.3, "
                "noise 1e-2, y = L z",
        "config": {"workload": "GP likelihood + gradient through torch.autograd (Cholesky and solve adjoints)",
                   "n": n, "tile": gp_b, "lengthscale": ell0, "noise": noise, "tuned_factor": tuned},
        "roofline": {"bound": "mfma", "achieved": round(flops / t_dev / 1e12, 3), "peak": PEAK["fp64"],
                     "unit": "TFLOP/s", "frac": round(flops / t_dev / 1e12 / PEAK["fp64"], 4), "traffic": None,
                     "basis": "n^3 flops (factor n^3/3, adjoint 2 n^3/3)"},
        "ms_per_step_note": "wall time per step: torch.autograd on the host around the device passes; K is "
                            "built on the device, only the points, y and scalars cross PCIe",
        "dnll_dl": g, "dnll_dl_central_difference": fd,
        "rel_err_vs_central_difference": abs(g - fd) / abs(fd),
        "rel_err_vs_cpu_closed_form": (abs(g - dell_ref) / abs(dell_ref)) if dell_ref is not None else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_hyper(args):
    """The hypermatrix factor (ompss_{d,s}chol_hrl, chol_hrlmain.c:16-37) in
    device memory: 2080 separate 512 x 512 tiles at n = 32768, timed beside the
    column-major factor of the same matrix (same plan and unit bodies)."""
    import numpy as np
    import linompss_amd as lx
    n, b = args.n, args.b
    if n % b:
        print("--n must be a multiple of --b for the hypermatrix workload", file=sys.stderr)
        sys.exit(2)
    mt = n // b
    flops = n ** 3 / 3.0
    log(f"generating A (n={n}, cond2~{args.kappa:g}, seed {args.seed}); {mt * (mt + 1) // 2} tiles of {b} x {b}")
    A, alpha, lmax = lx.gen_gauss_spd(n, args.seed, args.kappa)
    hA = A.to_host()
    res = {}
    for dt, name, up in ((np.float64, "fp64", 0), (np.float32, "fp32_split_bf16", 1)):
        hA_t = hA.astype(dt, order="F") if dt != np.float64 else hA
        pristine = {(i, j): lx.DeviceMatrix.from_host(np.asfortranarray(hA_t[i * b:(i + 1) * b, j * b:(j + 1) * b]))
                    for j in range(mt) for i in range(j, mt)}
        work = {k: lx.DeviceMatrix(b, b, dt) for k in pristine}
        opts = lx.ExecOpts(1, 128, 128, 0, 30.0, 0, up)
        t_tiles = []
        for k in range(args.warmup + args.steps):
            for key, m in pristine.items():
                lx.copy_dev(work[key], m)
            lx.lib.linompss_dev_sync()
            info = lx.chol_tiles_dev(work, mt, b, "rl", opts)
            if k >= args.warmup:
                t_tiles.append(lx.last_exec_stats().exec_ms * 1e-3)
            if info != 0:
                raise RuntimeError(f"tiled factor info {info}")
        for m in list(pristine.values()) + list(work.values()):
            m.free()
        # the column-major factor of the same matrix, same options
        src = lx.DeviceMatrix.from_host(hA_t)
        d = lx.DeviceMatrix(n, n, dt)
        t_col = []
        for k in range(args.warmup + args.steps):
            lx.copy_dev(d, src)
            lx.lib.linompss_dev_sync()
            info = lx.chol_dev(d, b, "rl", opts)
            if k >= args.warmup:
                t_col.append(lx.last_exec_stats().exec_ms * 1e-3)
            if info != 0:
                raise RuntimeError(f"column-major factor info {info}")
        src.free()
        d.free()
        tt, tc = float(np.median(t_tiles)), float(np.median(t_col))
        res[name] = {"tiles_ms": round(tt * 1e3, 3), "colmajor_ms": round(tc * 1e3, 3),
                     "tiles_tflops": round(flops / tt / 1e12, 3), "tiles_vs_colmajor": round(tc / tt, 4)}
        log(f"{name}: tiles {tt * 1e3:.1f} ms ({flops / tt / 1e12:.1f} TF), column-major {tc * 1e3:.1f} ms")
    A.free()
    t64 = res["fp64"]["tiles_ms"] * 1e-3
    out = {
        "metric": f"hypermatrix (tile-contiguous) Cholesky GFLOP/s at n={n}, b={b} (FP64, ompss_dchol_hrl on device)",
        "value": round(flops / t64 / 1e9, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t64 * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": f"synthetic: A = G G^T/n + alpha I, G ~ N(0,1) (Philox4x32-10, seed {args.seed}), as {mt * (mt + 1) // 2} separate device tiles",
        "config": {"workload": "hypermatrix factor, tiles as separate device buffers (chol_hrlmain.c:16-37)",
                   "n": n, "b": b, "mt": mt},
        "roofline": {"bound": "mfma", "achieved": round(flops / t64 / 1e12, 3), "peak": PEAK["fp64"],
                     "unit": "TFLOP/s", "frac": round(flops / t64 / 1e12 / PEAK["fp64"], 4), "traffic": None},
        "results": res,
    }
    print(json.dumps(out))


def bench_blas3(args):
    """The standalone BLAS-3 entry points on device memory (src/gemm/ompss_mm.h,
    src/syrk/ompss_syrk.c, src/trsm/ompss_trsm.c): GEMM and SYRK at
    n = k = 8192, left lower TRSM with 8192 x 512 right-hand sides, FP64 and FP32."""
    import ctypes
    import numpy as np
    import linompss_amd as lx
    n, nr = 8192, 512
    rng = np.random.default_rng(5)
    res = {}
    for dt, name, dtype in ((np.float64, "fp64", 0), (np.float32, "fp32", 1)):
        A = lx.DeviceMatrix.from_host(np.asfortranarray(rng.random((n, n)).astype(dt)))
        B = lx.DeviceMatrix.from_host(np.asfortranarray(rng.random((n, n)).astype(dt)))
        C = lx.DeviceMatrix.from_host(np.asfortranarray(rng.random((n, n)).astype(dt)))
        Lh = np.asfortranarray((np.tril(rng.random((n, n))) / n + np.eye(n)).astype(dt))
        L = lx.DeviceMatrix.from_host(Lh)
        X = lx.DeviceMatrix.from_host(np.asfortranarray(rng.random((n, nr)).astype(dt)))

        def timed(fn):
            ts = []
            for k in range(args.warmup + args.steps):
                lx.lib.linompss_dev_sync()
                t0 = time.perf_counter()
                rc = fn()
                lx.lib.linompss_dev_sync()
                if rc != 0:
                    raise RuntimeError(f"rc {rc}")
                if k >= args.warmup:
                    ts.append(time.perf_counter() - t0)
            return float(np.median(ts))
        t_gemm = timed(lambda: lx.lib.linompss_gemm_dev(dtype, b"N", b"T", n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.5,
                                                        C.ptr, n, None))
        t_syrk = timed(lambda: lx.lib.linompss_syrk_dev(dtype, b"L", b"N", n, n, 1.0, A.ptr, n, 0.5, C.ptr, n, None))
        t_trsm = timed(lambda: lx.lib.linompss_trsm_blas_dev(dtype, b"L", b"L", b"N", b"N", n, nr, 1.0, L.ptr, n,
                                                             X.ptr, n, None))
        peak = PEAK[name]
        res[name] = {
            "gemm_ms": round(t_gemm * 1e3, 3), "gemm_tflops": round(2.0 * n ** 3 / t_gemm / 1e12, 2),
            "gemm_frac": round(2.0 * n ** 3 / t_gemm / 1e12 / peak, 4),
            "syrk_ms": round(t_syrk * 1e3, 3), "syrk_tflops": round(1.0 * n ** 3 / t_syrk / 1e12, 2),
            "trsm_ms": round(t_trsm * 1e3, 3), "trsm_tflops": round(1.0 * n * n * nr / t_trsm / 1e12, 2),
        }
        log(f"{name}: gemm {t_gemm * 1e3:.2f} ms ({res[name]['gemm_tflops']} TF), syrk {t_syrk * 1e3:.2f} ms, "
            f"trsm {t_trsm * 1e3:.2f} ms")
        for m in (A, B, C, L, X):
            m.free()
    g64 = res["fp64"]["gemm_ms"] * 1e-3
    out = {
        "metric": "standalone blocked BLAS-3 on device: DGEMM TFLOP/s at m = n = k = 8192 (NT)",
        "value": res["fp64"]["gemm_tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(g64 * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: U[0,1) matrices (numpy PCG64 seed 5); TRSM on tril(U[0,1))/n + I",
        "config": {"workload": "ompss_{d,s}gemm / syrk / trsm entry points on device memory", "n": n,
                   "trsm_nrhs": nr},
        "roofline": {"bound": "mfma", "achieved": res["fp64"]["gemm_tflops"], "peak": PEAK["fp64"],
                     "unit": "TFLOP/s", "frac": res["fp64"]["gemm_frac"], "traffic": None},
        "results": res,
    }
    print(json.dumps(out))


def traffic_from(fetch_csv, write_csv, out, n, update_prec):
    """HBM traffic of the FP32 factor launch from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what the
    same counter reads for the device copy of A (lx_copy_kernel: n*n*8 bytes
    read and written) in the same run."""
    import csv
    known_kib = n * n * 8 / 1024.0

    def read(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copy = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"].startswith("lx_copy_kernel")]
        # the column-major FP32 factor instantiation ("<float, 0>" or "<float, 0, false>")
        fac = [float(r["Counter_Value"]) for r in rows
               if r["Kernel_Name"].startswith("void lx_exec_kernel<float, 0") and "true" not in r["Kernel_Name"]]
        if not copy or not fac:
            raise SystemExit(f"{path}: no lx_copy_kernel / FP32 factor dispatch with {counter}")
        return max(copy), sum(fac) / len(fac), len(fac)

    fc, ff, nf = read(fetch_csv, "FETCH_SIZE")
    wc, wf, nw = read(write_csv, "WRITE_SIZE")
    fs, ws = known_kib / fc, known_kib / wc
    res = {
        "kernel": "lx_exec_kernel<float, 0> (FP32 tiled factor, one persistent launch)",
        "workload": f"n={n}, b=512, FP32 factor inside the mixed-precision solve, trailing updates on "
                    f"{['FP32', 'split-BF16', 'split-FP16'][update_prec]} matrix cores",
        "update_prec": update_prec,
        "copy_calibration": {"bytes": n * n * 8, "fetch_kib": fc, "write_kib": wc,
                             "fetch_scale": fs, "write_scale": ws},
        "factor_fetch_kib": ff, "factor_write_kib": wf, "dispatches": [nf, nw],
        "hbm_read_bytes": ff * fs * 1024, "hbm_write_bytes": wf * ws * 1024,
        "hbm_bytes_per_launch": round((ff * fs + wf * ws) * 1024),
        "commands": [
            "rocprofv3 --pmc FETCH_SIZE -d <dir> -o run --output-format csv -- python3 bench.py --steps 1 "
            f"--warmup 0 --fp64-steps 1 --no-cpu --no-tune --update-prec {update_prec}",
            "rocprofv3 --pmc WRITE_SIZE -d <dir> -o run --output-format csv -- python3 bench.py --steps 1 "
            f"--warmup 0 --fp64-steps 1 --no-cpu --no-tune --update-prec {update_prec}",
            "python3 bench.py --traffic-from <fetch>/run_counter_collection.csv "
            f"<write>/run_counter_collection.csv --update-prec {update_prec} "
            "--traffic-out profiles/traffic_fp32_factor_n32768.json"],
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
    ap.add_argument("--fp64-b", type=int, default=0, help="tile of the plain FP64 factor (0: same as --b)")
    ap.add_argument("--cpu-n", type=int, default=32768)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--update-prec", type=int, default=2, choices=(0, 1, 2),
                    help="FP32 factor's trailing SYRK/GEMM: 0 FP32 matrix cores, 1 split-BF16 matrix cores "
                         "(six products), 2 split-FP16 matrix cores (three products, equilibrated A)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="derive the factor launch's HBM traffic from two rocprofv3 --pmc CSVs and exit")
    ap.add_argument("--traffic-out", default=str(ROOT / "profiles" / "traffic_fp32_factor_n32768.json"))
    ap.add_argument("--workload", default="mixed", choices=("mixed", "batched", "gp", "hyper", "blas3"),
                    help="mixed: the n=32768 refinement solve (the headline line); batched: 4096 "
                         "independent systems of order uniform in [32, 512] in one persistent launch; "
                         "gp: Gaussian-process likelihood forward + backward (Cholesky adjoint) at 8192 points; "
                         "hyper: the hypermatrix factor (separate device tiles) beside the column-major one; "
                         "blas3: the standalone GEMM / SYRK / TRSM entry points at n = 8192")
    ap.add_argument("--gp-n", type=int, default=8192)
    ap.add_argument("--gp-b", type=int, default=256, help="tile of the GP workload's FP64 factor")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--no-tune", action="store_true",
                    help="run the factors with the library's default unit shapes (no autotuning)")
    ap.add_argument("--no-tile-search", action="store_true",
                    help="skip the extra runs at the tile size the autotuner finds fastest")
    args = ap.parse_args()
    if args.workload in ("batched", "gp", "hyper", "blas3"):
        if args.gpus != 1:
            print("bench.py measures one GPU. Run with --gpus 1.", file=sys.stderr)
            sys.exit(2)
        {"batched": bench_batched, "gp": bench_gp, "hyper": bench_hyper, "blas3": bench_blas3}[args.workload](args)
        return
    if args.traffic_from:
        traffic_from(args.traffic_from[0], args.traffic_from[1], args.traffic_out, args.n, args.update_prec)
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

    # ---- autotuning (outside every timed region) ----
    # the unit shapes of both factors at the configuration's tile size
    fp64_b = args.fp64_b or b
    tuned = {}
    if not args.no_tune:
        for key, dt, bb, up in (("mixed_factor", np.float32, b, args.update_prec if args.update_prec else 0),
                                ("fp64_factor", np.float64, fp64_b, 0)):
            t0 = time.perf_counter()
            best, tried = lx.tune_factor(n, bb, dt, "rl", up, reps=2)
            tuned[key] = dict(best.as_dict(), candidates=len(tried), tune_s=round(time.perf_counter() - t0, 1))
            log(f"tuned {key}: {tuned[key]}")

    def run_mixed(bb, steps, warmup, tag):
        ir_opts = lx.IrOpts(30, 1e-14, 0.5, args.update_prec)
        for w in range(warmup):
            st = lx.dposv_ir_dev(A, B, X, bb, ir_opts)
            log(f"{tag} warmup {w}: sweeps={st.sweeps} fallback={st.fallback} berr={st.berr:.2e} "
                f"factor={st.t_factor_s * 1e3:.1f} ms total={st.t_total_s * 1e3:.1f} ms")
        lx.lib.linompss_dev_sync()
        r = {"t_fac": [], "t_launch": [], "sweeps": [], "berrs": [], "fallbacks": []}
        t0 = time.perf_counter()
        for k in range(steps):
            st = lx.dposv_ir_dev(A, B, X, bb, ir_opts)
            r["t_fac"].append(st.t_factor_s)
            # HIP events around the FP32 factor launch, on its stream (the last
            # executor launch of the solve; the sweeps and residuals use other kernels)
            r["t_launch"].append(lx.last_exec_stats().exec_ms * 1e-3)
            r["sweeps"].append(st.sweeps)
            r["berrs"].append(st.berr)
            r["fallbacks"].append(st.fallback)
        lx.lib.linompss_dev_sync()
        r["t"] = (time.perf_counter() - t0) / steps
        r["berr"] = lx.dberr_dev(A, X, B)
        log(f"{tag}: {r['t'] * 1e3:.1f} ms/solve, factor {np.mean(r['t_fac']) * 1e3:.1f} ms, sweeps {r['sweeps']}, "
            f"berr {r['berr']:.2e}, fallback {r['fallbacks']}")
        return r

    def run_fp64(bb, steps, tag):
        A2 = lx.DeviceMatrix(n, n)
        X2 = lx.DeviceMatrix(n, nr)
        t64, launch = [], []
        for k in range(max(1, steps)):
            lx.copy_dev(A2, A)
            lx.copy_dev(X2, B)
            lx.lib.linompss_dev_sync()
            t0 = time.perf_counter()
            info = lx.posv_dev(A2, X2, bb)
            lx.lib.linompss_dev_sync()
            t64.append(time.perf_counter() - t0)
            st64 = lx.last_exec_stats()
            launch.append(st64.exec_ms)
            log(f"{tag} {k}: {t64[-1] * 1e3:.1f} ms (launch {st64.exec_ms:.1f} ms, units {st64.units}) info={info}")
        berr = lx.dberr_dev(A, X2, B)
        A2.free()
        X2.free()
        return min(t64), min(launch), berr

    # ---- mixed-precision refinement (the metric) ----
    mx = run_mixed(b, args.steps, args.warmup, "mixed")
    t_ir, berr_ir = mx["t"], mx["berr"]
    t_fac, t_launch, sweeps, fallbacks = mx["t_fac"], mx["t_launch"], mx["sweeps"], mx["fallbacks"]

    # ---- plain FP64 factor + solve (one factor launch + the sweeps) ----
    t_fp64, t_fp64_launch, berr64 = run_fp64(fp64_b, args.fp64_steps, "fp64 posv")

    # ---- both paths again at the tile size the autotuner finds fastest ----
    tile_search = None
    if not args.no_tune and not args.no_tile_search:
        ts = {}
        for key, dt, up in (("mixed", np.float32, args.update_prec if args.update_prec else 0),
                            ("fp64", np.float64, 0)):
            t0 = time.perf_counter()
            best, tried = lx.tune_factor(n, 0, dt, "rl", up, reps=2)
            ts[key] = {"config": best.as_dict(), "candidates": len(tried), "tune_s": round(time.perf_counter() - t0, 1)}
            log(f"tile search {key}: {ts[key]}")
        m2 = run_mixed(ts["mixed"]["config"]["b"], max(3, args.steps // 2), 1, "mixed (tuned tile)")
        f2, f2_launch, f2_berr = run_fp64(ts["fp64"]["config"]["b"], args.fp64_steps, "fp64 posv (tuned tile)")
        ts["mixed"].update({"ms": round(m2["t"] * 1e3, 3), "gflops": round(flops / m2["t"] / 1e9, 2),
                            "factor_launch_ms": round(float(np.mean(m2["t_launch"])) * 1e3, 3),
                            "berr": m2["berr"], "sweeps": int(np.max(m2["sweeps"])),
                            "fallback": int(np.max(m2["fallbacks"]))})
        ts["fp64"].update({"ms": round(f2 * 1e3, 3), "gflops": round(flops / f2 / 1e9, 2),
                           "factor_launch_ms": round(f2_launch, 3),
                           "factor_frac_of_fp64_peak": round(flops / (f2_launch * 1e-3) / 1e12 / PEAK["fp64"], 4),
                           "berr": f2_berr})
        ts["speedup_vs_fp64"] = round(f2 / m2["t"], 3)
        tile_search = ts

    t_kernel = float(np.mean(t_launch))
    achieved = flops / t_kernel / 1e12
    if args.update_prec == 1:
        # six BF16 products per FP32 multiply-add of the trailing updates
        peak, basis = PEAK["bf16"] / 6.0, "bf16 dense 2500 TFLOP/s / 6 split products"
    elif args.update_prec == 2:
        # three FP16 products per FP32 multiply-add (FP16 and BF16 share the dense rate)
        peak, basis = PEAK["bf16"] / 3.0, "fp16 dense 2500 TFLOP/s / 3 split products"
    else:
        peak, basis = PEAK["fp32"], "fp32 dense 157.3 TFLOP/s"
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": round(peak, 1),
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4),
        "traffic": None,
        "peak_basis": basis,
        "frac_of_fp32_peak": round(achieved / PEAK["fp32"], 4),
    }
    tf = ROOT / "profiles" / "traffic_fp32_factor_n32768.json"
    if tf.exists() and (n, b) == (32768, 512):
        try:
            rec = json.loads(tf.read_text())
            if rec.get("update_prec", 1) == args.update_prec:  # the workload the counters were read on
                roofline["traffic"] = rec.get("hbm_bytes_per_launch")
        except Exception:
            pass

    cpu = None
    if not args.no_cpu:
        threads = host_threads(args.cpu_threads)
        log(f"CPU baseline: n={args.cpu_n} on {threads} threads")
        try:
            cpu = cpu_baseline(args.cpu_n, b, threads)
            log(f"CPU baseline {cpu['value']} GFLOP/s")
        except Exception as e:  # the oracle needs the host BLAS it was built against
            log(f"CPU baseline unavailable: {e}")

    out = {
        "metric": METRIC if n == 32768 else METRIC.replace("n=32768", f"n={n}"),
        "value": round(flops / t_ir / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_ir * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ["f32 factor + f64 refinement",
                  "f32 factor (trailing SYRK/GEMM on bf16 matrix cores, split operands hi+mid+lo, f32 "
                  "accumulation) + f64 refinement",
                  "f32 factor of the equilibrated matrix (trailing SYRK/GEMM on fp16 matrix cores, split "
                  "operands hi+lo of x*2^12, f32 accumulation) + f64 refinement"][args.update_prec],
        "data": "synthetic: A = G G^T/n + alpha I, G ~ N(0,1) (Philox4x32-10, seed %d), B ~ U[-1,1)" % args.seed,
        "config": {"workload": "FP32-factor + FP64 iterative refinement, SPD solve to backward error 1e-14",
                   "update_prec": ["fp32", "split-bf16", "split-fp16"][args.update_prec],
                   "n": n, "b": b, "cond2": args.kappa, "nrhs": nr, "alpha": alpha, "lambda_max": lmax},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "berr": berr_ir,
        "sweeps": int(np.max(sweeps)),
        "fallback": int(np.max(fallbacks)),
        "factor_ms": round(float(np.mean(t_fac)) * 1e3, 3),
        "factor_launch_ms": round(t_kernel * 1e3, 3),
        "fp64_posv_ms": round(t_fp64 * 1e3, 3),
        "fp64_b": fp64_b,
        "fp64_posv_gflops": round(flops / t_fp64 / 1e9, 2),
        "fp64_factor_launch_ms": round(t_fp64_launch, 3),
        "fp64_factor_frac_of_fp64_peak": round(flops / (t_fp64_launch * 1e-3) / 1e12 / PEAK["fp64"], 4),
        "fp64_berr": berr64,
        "speedup_vs_fp64": round(t_fp64 / t_ir, 3),
        "tuned": tuned or None,
        "tile_search": tile_search,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
