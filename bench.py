#!/usr/bin/env python3
"""Headline benchmark: FP64-accurate dense solve on one MI355X.

Workload (the configuration the headline metric is quoted on): n = 16384,
A = U diag(sigma) V^T with U, V products of 64 Householder reflectors
(seeded normal vectors), sigma geometrically spaced from 1 to 1/kappa,
kappa = 1e6, one standard-normal right-hand side.  A "step" is one complete
FP64-accurate solve with inputs resident in HBM:
  FP32-IR : convert A to FP32, blocked LU on FP32 MFMA, FP64 iterative
            refinement to normwise backward error <= 1e-15 * n
  FP64    : plain blocked FP64 LU with partial pivoting + triangular solves
value = FP64-equivalent flop count (2/3 n^3 - n^2/2 + 5n/6 + 2 n^2) divided
by the FP32-IR step time (TFLOP/s).  The plain FP64 solve is timed beside
it on the same A and b.

Also reported: backward and forward errors, refinement steps, the dominant
kernel's roofline (HIP-event timing of every launch of that kernel class
during one extra, instrumented solve), and the CPU oracle (the reference's
algorithm restated in C, one core) on a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rust-la_amd"))

PEAKS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0, "hbm_gbs": 8000.0}


def lu_solve_flop(n, nrhs=1):
    return (2.0 / 3.0) * n ** 3 - 0.5 * n ** 2 + (5.0 / 6.0) * n + 2.0 * n ** 2 * nrhs


def cpu_baseline(n_sample):
    """The CPU oracle (C restatement of lu.rs:104-278, -O2, no FMA, one
    thread) timed on an n_sample x n_sample uniform system."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import numpy as np
    a = oracle.uniform(7, 0, (n_sample, n_sample))
    b = oracle.uniform(7, 2, (n_sample, 1))
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    t0 = time.perf_counter()
    lu, piv, _ = oracle.lu_new(a)
    oracle.lu_solve(lu, piv, b)
    dt = time.perf_counter() - t0
    return {"value": lu_solve_flop(n_sample) / dt / 1e12, "unit": "TFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"LU with partial pivoting + solve, n={n_sample}, 1 RHS, uniform[-1,1] "
                      f"({dt:.1f} s on one core; C restatement of lu.rs:104-278)",
            "seconds": dt}


def batched_cholesky(ctx, steps, warmup, count=20000, seed=20000):
    """The batched uneven Cholesky workload: `count` SPD problems, n_i uniform
    in [8, 256], A_i = G_i G_i^T / n_i + I, one normal RHS each; potrf plus
    both triangular solves in one graph-replayed persistent launch."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    sizes = rng.integers(8, 257, size=count).astype(np.int32)
    n64 = sizes.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 ** 2)[:-1]])
    offB = np.concatenate([[0], np.cumsum(n64)[:-1]])
    dn, dOA, dOB = (torch.from_numpy(x).cuda() for x in (sizes, offA, offB))
    A = torch.empty(int((n64 ** 2).sum()), dtype=torch.float64, device="cuda")
    G = torch.empty_like(A)
    ctx.gen_spd_batch_dev(dn, dOA, 4, 1000, A, G)
    del G
    B = torch.empty(int(n64.sum()), dtype=torch.float64, device="cuda")
    ctx.fill_normal_dev(4, 7, B)
    L, X = torch.empty_like(A), torch.empty_like(B)
    st = torch.empty(count, dtype=torch.int32, device="cuda")
    ctx.autotune(0, 8, 256)
    plan = ctx.batch_plan(0, sizes)
    for _ in range(max(1, warmup)):
        ctx.potrf_potrs_batched_dev(plan, dOA, dOB, A, L, B, X, st)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        ctx.potrf_potrs_batched_dev(plan, dOA, dOB, A, L, B, X, st)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    failed = int((st != 0).sum().item())
    nbytes = float((8 * (2 * n64 ** 2 + 2 * n64)).sum())
    flop = float((n64 ** 3 / 3 + n64 ** 2 / 2 + n64 / 6 + 4 * n64 ** 2).sum())
    # spot-check a spread sample of backward errors on the host
    a, b, x = A.cpu().numpy(), B.cpu().numpy(), X.cpu().numpy()
    worst = 0.0
    for i in range(0, count, 97):
        n = int(sizes[i])
        ai = a[offA[i]:offA[i] + n * n].reshape(n, n)
        xi, bi = x[offB[i]:offB[i] + n], b[offB[i]:offB[i] + n]
        eta = np.abs(bi - ai @ xi).max() / (np.abs(ai).sum(1).max() * np.abs(xi).max()
                                            + np.abs(bi).max())
        worst = eta if not np.isfinite(eta) else max(worst, eta)
    desc = plan.describe()
    plan.close()
    return {"count": count, "n_range": [8, 256], "ms": ms, "solves_per_s": count / (ms * 1e-3),
            "hbm_gbs": nbytes / (ms * 1e-3) / 1e9,
            "frac_of_hbm_roofline": nbytes / (ms * 1e-3) / 1e9 / PEAKS["hbm_gbs"],
            "fp64_tflops": flop / (ms * 1e-3) / 1e12, "failed": failed,
            "max_backward_error_sampled": worst, "plan": desc}


def batched_lu_ir(ctx, steps, warmup, count=4096, seed=4096):
    """The batched uneven LU-IR workload: `count` general problems, n
    log-uniform in [32, 1024], A_i = U diag(sigma) V^T with kappa = 1e3
    (64 Householder reflectors each), 1-8 normal right-hand sides; FP16
    factorisation refined in FP64 to eta <= 1e-14, one graph-replayed
    persistent launch.  The batched plain FP64 LU runs beside it."""
    import numpy as np
    import torch
    import la_hip
    rng = np.random.default_rng(seed)
    sizes = np.round(32 * 32 ** rng.random(count)).astype(np.int32)
    nrhs = rng.integers(1, 9, size=count).astype(np.int32)
    n64 = sizes.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 ** 2)[:-1]])
    offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]])
    dn, dr, dOA, dOB = (torch.from_numpy(x).cuda() for x in (sizes, nrhs, offA, offB))
    A = torch.empty(int((n64 ** 2).sum()), dtype=torch.float64, device="cuda")
    B = torch.empty(int((n64 * nrhs).sum()), dtype=torch.float64, device="cuda")
    Xs = torch.empty_like(B)
    ctx.gen_svd_batch_dev(dn, dr, dOA, dOB, 1e3, 64, 11, 0, A, B, Xs)
    flop = float((2.0 / 3.0 * n64 ** 3).sum())
    res = {"count": count, "n_range": [32, 1024], "kappa": 1e3,
           "fp64_equiv_flop": flop}
    for name, prec in (("fp16_ir", la_hip.PREC_FP16), ("fp64_lu", la_hip.PREC_FP64)):
        plan = ctx.batch_plan(1, sizes, nrhs, prec)
        X = torch.empty_like(B)
        st = torch.empty(count, dtype=torch.int32, device="cuda")
        it = torch.empty(count, dtype=torch.int32, device="cuda")
        be = torch.empty(count, dtype=torch.float64, device="cuda")
        for _ in range(max(1, warmup)):
            ctx.gesv_ir_batched_dev(plan, dOA, dOB, A, B, X, 1e-14, 30, st, it, be)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            ctx.gesv_ir_batched_dev(plan, dOA, dOB, A, B, X, 1e-14, 30, st, it, be)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        fwd = ((X - Xs).abs().max() / Xs.abs().max()).item()
        res[name] = {"ms": ms, "solves_per_s": count / (ms * 1e-3),
                     "fp64_equiv_tflops": flop / (ms * 1e-3) / 1e12,
                     "failed": int((st != 0).sum().item()),
                     "max_backward_error": be.max().item(),
                     "mean_refinement_steps": it.double().mean().item(),
                     "max_refinement_steps": int(it.max().item()),
                     "max_forward_error_vs_xstar": fwd, "plan": plan.describe()}
        if name == "fp16_ir":
            X16 = X.clone()
        else:
            res["max_diff_fp16ir_vs_fp64lu"] = ((X16 - X).abs().max() / X.abs().max()).item()
        plan.close()
    return res


# measured HBM bytes per launch (read + write) by kernel class, dense solve n = 16384
MEASURED_TRAFFIC = {"gemm32": 91.87e6}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--cpu-n", type=int, default=1536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the solver and its "
              "workloads are single-device.", file=sys.stderr)
        sys.exit(2)

    import torch
    import la_hip
    ctx = la_hip.Context(0)
    ctx.use_torch_stream()
    n, kappa = args.n, args.kappa
    dev = "cuda"
    A = torch.empty((n, n), dtype=torch.float64, device=dev)
    b = torch.empty((n, 1), dtype=torch.float64, device=dev)
    xs = torch.empty_like(b)
    ctx.gen_svd_matrix_dev(n, kappa, 64, 2025, A, b=b, xstar=xs)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    torch.cuda.synchronize()

    # ---------------- FP32-IR (the measured path): butterfly-transformed FP32
    # LU without pivoting, FP64 refinement on A (falls back to the pivoted
    # factor if refinement does not converge; the report says which ran)
    def ir_step(prec=la_hip.PREC_FP32 | la_hip.IR_RBT):
        return ctx.gesv_ir_dev(A, b, X, prec, tol, 50)

    for _ in range(args.warmup):
        ir_step()
    torch.cuda.synchronize()
    times, reps = [], []
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rep = ir_step()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
        reps.append(rep)
    wall = time.perf_counter() - t_all0
    ms_ir = sum(times) / len(times)
    rep = reps[-1]
    fwd_ir = ((X - xs).abs().max() / xs.abs().max()).item()
    eta_ir = ctx.backward_error_dev(A, X, b)

    # ---------------- FP32-IR with the partial-pivoting FP32 factor
    Xp = X.clone()
    ir_step(la_hip.PREC_FP32)
    torch.cuda.synchronize()
    tpv = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rep_pv = ir_step(la_hip.PREC_FP32)
        e1.record()
        torch.cuda.synchronize()
        tpv.append(e0.elapsed_time(e1))
    ms_pv = sum(tpv) / len(tpv)
    eta_pv = ctx.backward_error_dev(A, X, b)
    fwd_pv = ((X - xs).abs().max() / xs.abs().max()).item()
    X.copy_(Xp)

    # ---------------- plain FP64 LU + solve on the same system
    W = torch.empty_like(A)
    piv = torch.empty(n, dtype=torch.int64, device=dev)
    info = torch.zeros(4, dtype=torch.int32, device=dev)
    st = torch.zeros(1, dtype=torch.int32, device=dev)
    X64 = torch.empty_like(b)
    t64 = []
    for it in range(args.warmup + args.steps):
        W.copy_(A)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ctx.getrf_dev(W, piv, info)
        ctx.getrs_dev(W, piv, b, X64, st)
        e1.record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            t64.append(e0.elapsed_time(e1))
    ms_64 = sum(t64) / len(t64)
    eta_64 = ctx.backward_error_dev(A, X64, b)
    fwd_64 = ((X64 - xs).abs().max() / xs.abs().max()).item()

    # ---------------- instrumented solve: per-kernel-class HIP-event timing
    ctx.reset_kernel_stats()
    ctx.set_profiling(True)
    ir_step()
    torch.cuda.synchronize()
    ctx.set_profiling(False)
    stats = ctx.kernel_stats()
    dom = max(stats.items(), key=lambda kv: kv[1][1])
    name, (calls, tot_ms, flop, nbytes) = dom
    avg_ms = tot_ms / calls
    if name.startswith("gemm"):
        peak = {"gemm64": PEAKS["fp64"], "gemm32": PEAKS["fp32"], "gemm16": PEAKS["fp16"]}[name]
        achieved = flop / calls / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s"}
    else:
        achieved = nbytes / calls / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAKS["hbm_gbs"], "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    # HBM bytes per launch of the dominant kernel on this workload (n = 16384), from the two
    # rocprofv3 counter passes FETCH_SIZE (x 2.000) and WRITE_SIZE (x 1.000), the factors
    # calibrated on a 2 GiB copy in the same runs: profiles/gemm32_hbm_traffic_n16384.md
    roof["traffic"] = MEASURED_TRAFFIC.get(name) if n == 16384 else None
    roof["kernel"] = name
    roof["avg_launch_ms"] = avg_ms
    roof["launches_per_step"] = calls

    flop = lu_solve_flop(n)
    out = {
        "metric": "FP64-accurate solve TFLOP/s: FP32-IR vs plain FP64 LU, n=16384; batched solves/s",
        "value": flop / (ms_ir * 1e-3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_ir,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-factor+fp64-refine",
        "data": "synthetic: seeded Householder-product A = U diag(sigma) V^T, normal b",
        "config": {"workload": f"dense solve n={n}, kappa={kappa:g}, 1 RHS, FP32 LU (random "
                               f"butterfly, no pivoting) + FP64 IR to eta<=1e-15*n",
                   "n": n, "kappa": kappa, "nrhs": 1},
        "fp32_ir": {"ms": ms_ir, "tflops": flop / (ms_ir * 1e-3) / 1e12,
                    "factor": "random butterfly U^T A V, FP32 LU without pivoting",
                    "method": ["pivoted", "butterfly", "butterfly then pivoted"][rep.method],
                    "perturbed_pivots": rep.perturbed_pivots,
                    "iterations": rep.iterations, "backward_error": eta_ir,
                    "forward_error_vs_xstar": fwd_ir, "factor_ms": rep.factor_ms,
                    "refine_ms": rep.refine_ms, "status": rep.status},
        "fp32_ir_pivoted": {"ms": ms_pv, "tflops": flop / (ms_pv * 1e-3) / 1e12,
                            "factor": "FP32 LU with partial pivoting",
                            "iterations": rep_pv.iterations, "backward_error": eta_pv,
                            "forward_error_vs_xstar": fwd_pv, "factor_ms": rep_pv.factor_ms,
                            "refine_ms": rep_pv.refine_ms, "status": rep_pv.status},
        "fp64_lu": {"ms": ms_64, "tflops": flop / (ms_64 * 1e-3) / 1e12,
                    "backward_error": eta_64, "forward_error_vs_xstar": fwd_64,
                    "frac_of_fp64_peak": flop / (ms_64 * 1e-3) / 1e12 / PEAKS["fp64"]},
        "speedup_fp32ir_vs_fp64": ms_64 / ms_ir,
        "speedup_fp32ir_pivoted_vs_fp64": ms_64 / ms_pv,
        "roofline": roof,
        "kernel_breakdown_ms": {k: round(v[1], 3) for k, v in stats.items()},
        "wall_s_timed": wall,
    }
    if not args.no_batched:
        out["batched_cholesky"] = batched_cholesky(ctx, args.steps, args.warmup)
        out["batched_lu_ir"] = batched_lu_ir(ctx, args.steps, args.warmup)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
        out["cpu_baseline"]["host_nproc"] = os.cpu_count()
        try:
            model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                     if l.startswith("model name")][0]
            out["cpu_baseline"]["host_cpu"] = model
        except Exception:
            pass
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
