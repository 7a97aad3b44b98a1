#!/usr/bin/env python3
"""Headline benchmark: FP64-accurate dense solve on one MI355X.

Workload (the configuration the headline metric is quoted on): n = 16384,
A = U diag(sigma) V^T with U, V products of 64 Householder reflectors
(seeded normal vectors), sigma geometrically spaced from 1 to 1/kappa,
kappa = 1e6, one standard-normal right-hand side.  A "step" is one complete
FP64-accurate solve with its inputs resident in HBM: convert A to FP32,
factor it with partial pivoting (trailing updates on the FP32 MFMA), refine
in FP64 (residual, correction, update) to normwise backward error
<= 1e-15 n.  The timed steps cycle over five seeds
(five different systems of the family, generated before the timed region),
so the figure is not the behaviour of one matrix.

value = sum of FP64-equivalent flop ((2/3) n^3 - n^2/2 + 5n/6 + 2 n^2 per
solve) over the timed steps / their total time (TFLOP/s); it is null unless
every step converged.  The plain FP64 LU + solve (partial pivoting, the
reference's algorithm) is timed beside it on the same systems.

Also reported: the tournament-pivoted and butterfly FP32 factors on the
same systems, a kappa sweep
{1e2, 1e4, 1e6} x 5 seeds of every FP32 method (which factor served, steps,
eta, forward error, time), BF16-factor GMRES refinement at n = 16384, the
n = 8192 FP64 LU + 16-RHS solve and DGEMM 8192^3, both batched workloads,
the dominant kernel's roofline (HIP-event timing of every launch of that
kernel class during one instrumented solve, traffic per launch from the
committed counter summary), and the CPU oracle (the reference's algorithm
restated in C, one core) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rust-la_amd"))

PEAKS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0, "hbm_gbs": 8000.0}
SEEDS = [2025, 2026, 2027, 2028, 2029]
METRIC = "FP64-accurate solve TFLOP/s: FP32-IR vs plain FP64 LU, n=16384; batched solves/s"


def lu_solve_flop(n, nrhs=1):
    return (2.0 / 3.0) * n ** 3 - 0.5 * n ** 2 + (5.0 / 6.0) * n + 2.0 * n ** 2 * nrhs


def cpu_baseline(n_sample):
    """The CPU oracle (C restatement of lu.rs:104-278, -O2, no FMA, one
    thread) timed on an n_sample x n_sample uniform system."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
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


def cpu_sweep():
    """The CPU oracle (one core, -O2, no FMA) over the sizes of the path:
    LU + solve, Mul and Cholesky + solve at n = 256 .. 2048 (Mul to 1024),
    and a 2000-problem sample of the batched Cholesky workload (sizes drawn
    like the full 20000-problem batch).  About 10 s on the MI355X host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    out = {"unit": "GFLOP/s", "cores": 1, "kind": "port"}

    def clock(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0
    lu, mm, ch = {}, {}, {}
    for n in (256, 512, 1024, 2048):
        a = oracle.uniform(7, 0, (n, n))
        b = oracle.uniform(7, 2, (n, 1))
        dt = clock(lambda: oracle.lu_solve(*oracle.lu_new(a)[:2], b))
        lu[n] = {"s": dt, "gflops": lu_solve_flop(n) / dt / 1e9}
        spd = oracle.spd_matrix(n, 9, 1)
        dt = clock(lambda: oracle.chol_solve(oracle.chol_new(spd)[0], b))
        ch[n] = {"s": dt, "gflops": (n ** 3 / 3 + 2 * n ** 2) / dt / 1e9}
        if n <= 1024:
            dt = clock(lambda: oracle.mul(a, a))
            mm[n] = {"s": dt, "gflops": 2.0 * n ** 3 / dt / 1e9}
    out["lu_solve"], out["cholesky_solve"], out["mul"] = lu, ch, mm
    rng = np.random.default_rng(20000)
    sizes = rng.integers(8, 257, size=2000)
    mats = [oracle.spd_matrix(int(n), 4, i) for i, n in enumerate(sizes)]
    rhs = [oracle.normal(4, 7000 + i, (int(n), 1)) for i, n in enumerate(sizes)]
    dt = clock(lambda: [oracle.chol_solve(oracle.chol_new(m)[0], r) for m, r in zip(mats, rhs)])
    out["batched_cholesky_sample"] = {
        "problems": 2000, "n_range": [8, 256], "s": dt, "solves_per_s": 2000 / dt,
        "note": "sizes uniform in [8, 256] as in the 20000-problem workload; one core"}
    return out


def timed(torch, fn, reps=1):
    """Mean GPU time (ms) of fn() over reps calls, HIP events on the current stream."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    out = None
    for _ in range(reps):
        out = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, out


def method_name(rep):
    import la_hip
    return la_hip.IR_METHOD_NAMES[rep.method]


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
    ms, _ = timed(torch, lambda: ctx.potrf_potrs_batched_dev(plan, dOA, dOB, A, L, B, X, st),
                  steps)
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
    ctx.autotune(1, 32, 1024)
    res = {"count": count, "n_range": [32, 1024], "kappa": 1e3, "fp64_equiv_flop": flop}
    X16 = None
    for name, prec in (("fp16_ir", la_hip.PREC_FP16), ("fp32_ir", la_hip.PREC_FP32),
                       ("fp64_lu", la_hip.PREC_FP64)):
        plan = ctx.batch_plan(1, sizes, nrhs, prec)
        X = torch.empty_like(B)
        st = torch.empty(count, dtype=torch.int32, device="cuda")
        it = torch.empty(count, dtype=torch.int32, device="cuda")
        be = torch.empty(count, dtype=torch.float64, device="cuda")

        def run():
            ctx.gesv_ir_batched_dev(plan, dOA, dOB, A, B, X, 1e-14, 30, st, it, be)
        for _ in range(max(1, warmup)):
            run()
        ms, _ = timed(torch, run, steps)
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
        elif name == "fp64_lu":
            res["max_diff_fp16ir_vs_fp64lu"] = ((X16 - X).abs().max() / X.abs().max()).item()
        plan.close()
    return res


def qr_bench(ctx, torch):
    """Householder QR (QRDecomposition::new) of a standard-normal n x n
    matrix on the device, n = 4096 and 8192: ms and TFLOP/s at (4/3) n^3."""
    out = {}
    for n in (4096, 8192):
        A = torch.empty((n, n), dtype=torch.float64, device="cuda")
        ctx.fill_normal_dev(31, 0, A)
        W = torch.empty_like(A)
        rd = torch.empty(n, dtype=torch.float64, device="cuda")
        W.copy_(A)
        ctx.geqrf_dev(W, rd)
        W.copy_(A)
        ms, _ = timed(torch, lambda: ctx.geqrf_dev(W, rd))
        out[f"n{n}"] = {"ms": ms, "tflops": (4.0 / 3.0) * n ** 3 / (ms * 1e-3) / 1e12,
                        "frac_of_fp64_peak": (4.0 / 3.0) * n ** 3 / (ms * 1e-3) / 1e12 / PEAKS["fp64"]}
        del A, W
    return out


def cfg2(ctx, torch, reps):
    """n = 8192 plain FP64 LU + 16-RHS solve (standard normal A, B) and
    DGEMM 8192^3 on standard normal operands."""
    n = 8192
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    ctx.fill_normal_dev(2024, 0, A)
    B = torch.empty((n, 16), dtype=torch.float64, device="cuda")
    ctx.fill_normal_dev(2024, 2, B)
    W = torch.empty_like(A)
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    X = torch.empty_like(B)
    t = []
    for it in range(reps + 1):
        W.copy_(A)
        ms, _ = timed(torch, lambda: (ctx.getrf_dev(W, piv, info), ctx.getrs_dev(W, piv, B, X, st)))
        if it:
            t.append(ms)
    ms_lu = sum(t) / len(t)
    eta = ctx.backward_error_dev(A, X, B)
    flop_lu = lu_solve_flop(n, 16)
    C = torch.empty_like(A)
    ctx.gemm_dev(A, W, C)
    ms_gemm, _ = timed(torch, lambda: ctx.gemm_dev(A, W, C), reps)
    tf_lu = flop_lu / (ms_lu * 1e-3) / 1e12
    tf_gemm = 2.0 * n ** 3 / (ms_gemm * 1e-3) / 1e12
    return {"n": n, "fp64_lu_16rhs": {"ms": ms_lu, "tflops": tf_lu,
                                      "frac_of_fp64_peak": tf_lu / PEAKS["fp64"],
                                      "backward_error": eta, "status": int(info[1].item())},
            "dgemm_8192": {"ms": ms_gemm, "tflops": tf_gemm,
                           "frac_of_fp64_peak": tf_gemm / PEAKS["fp64"]}}


def next_rows(ctx, torch, reps):
    """The rows next to the path, timed on the device: Matrix::inverse at
    n = 8192 (LU, then the inverse by two MFMA TRSMs; (2/3 + 2) n^3 flop),
    Matrix<f32>::simd_mul's FP32 product at 8192^3, and dense Cholesky +
    one-RHS solve at n = 16384 ((1/3) n^3 + 2 n^2 flop)."""
    out = {}
    n = 8192
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    ctx.fill_normal_dev(2025, 0, A)
    W, X = torch.empty_like(A), torch.empty_like(A)
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    W.copy_(A)
    ctx.inverse_dev(W, piv, X, info)
    t = []
    for _ in range(reps):
        W.copy_(A)
        ms, _ = timed(torch, lambda: ctx.inverse_dev(W, piv, X, info))
        t.append(ms)
    ms = sum(t) / len(t)
    flop = (2.0 / 3.0 + 2.0) * n ** 3
    # residual of the inverse on a few columns: |A X[:, j] - e_j|
    cols = torch.arange(0, n, n // 8, device="cuda")
    R = A @ X[:, cols]
    R[cols, torch.arange(len(cols), device="cuda")] -= 1.0
    out["inverse_n8192"] = {"ms": ms, "tflops": flop / (ms * 1e-3) / 1e12,
                            "frac_of_fp64_peak": flop / (ms * 1e-3) / 1e12 / PEAKS["fp64"],
                            "max_abs_residual_sampled": float(R.abs().max().item()),
                            "status": int(info[1].item())}
    del W, X, A, R
    As = torch.randn(n, n, dtype=torch.float32, device="cuda")
    Bs = torch.randn(n, n, dtype=torch.float32, device="cuda")
    Cs = torch.empty_like(As)
    ctx.sgemm_dev(As, Bs, Cs)
    ms, _ = timed(torch, lambda: ctx.sgemm_dev(As, Bs, Cs), reps)
    tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
    out["sgemm_8192"] = {"ms": ms, "tflops": tf, "frac_of_fp32_peak": tf / PEAKS["fp32"]}
    del As, Bs, Cs
    n = 16384
    G = torch.empty((n, n), dtype=torch.float64, device="cuda")
    ctx.fill_normal_dev(2026, 0, G)
    S = torch.empty_like(G)
    ctx.gemm_dev(G, G.t().contiguous(), S, alpha=1.0 / n)
    S.diagonal().add_(1.0)
    del G
    L = torch.empty_like(S)
    b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    ctx.fill_normal_dev(2026, 1, b)
    x = torch.empty_like(b)
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    t = []
    for it in range(reps + 1):
        L.copy_(S)
        ms, _ = timed(torch, lambda: (ctx.potrf_dev(L, st), ctx.potrs_dev(L, b, x)))
        if it:
            t.append(ms)
    ms = sum(t) / len(t)
    flop = n ** 3 / 3.0 + 2.0 * n ** 2
    eta = ctx.backward_error_dev(S, x, b)
    out["cholesky_n16384"] = {"ms": ms, "tflops": flop / (ms * 1e-3) / 1e12,
                              "frac_of_fp64_peak": flop / (ms * 1e-3) / 1e12 / PEAKS["fp64"],
                              "backward_error": eta, "status": int(st.item())}
    return out


def load_traffic():
    """HBM bytes per launch by (kernel class, workload) from the committed
    counter summary (profiles/traffic_per_launch.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic_per_launch.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--method", default="pivoted", choices=["pivoted", "tournament", "rbt"])
    ap.add_argument("--cpu-n", type=int, default=2560)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--seeds", type=int, default=len(SEEDS),
                    help="systems the timed steps cycle over (profiling runs use fewer)")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the other FP32 factors (profiling runs)")
    ap.add_argument("--no-cfg2", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the solver and its "
              "workloads are single-device.", file=sys.stderr)
        sys.exit(2)

    import torch
    import la_hip
    n, kappa = args.n, args.kappa
    tol = 1e-15 * n
    flop = lu_solve_flop(n)
    out = {}
    with la_hip.Context(0) as ctx:
        ctx.use_torch_stream()
        opt = {"pivoted": 0, "tournament": la_hip.IR_TOURNAMENT, "rbt": la_hip.IR_RBT}
        head_prec = la_hip.PREC_FP32 | opt[args.method]

        # ---------------- the five systems of the headline, resident in HBM
        systems = []
        t_gen = []
        used_seeds = SEEDS[:max(1, args.seeds)]
        for s in used_seeds:
            A = torch.empty((n, n), dtype=torch.float64, device="cuda")
            b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
            xs = torch.empty_like(b)
            ms, _ = timed(torch, lambda: ctx.gen_svd_matrix_dev(n, kappa, 64, s, A, b=b, xstar=xs))
            t_gen.append(ms)
            systems.append((s, A, b, xs))
        X = torch.empty((n, 1), dtype=torch.float64, device="cuda")

        # ---------------- headline: FP32-IR, steps cycle over the systems
        for i in range(args.warmup):
            _, A, b, _ = systems[i % len(systems)]
            ctx.gesv_ir_dev(A, b, X, head_prec, tol, 50)
        torch.cuda.synchronize()
        steps, t_all0 = [], time.perf_counter()
        for i in range(args.steps):
            s, A, b, xs = systems[i % len(systems)]
            ms, rep = timed(torch, lambda: ctx.gesv_ir_dev(A, b, X, head_prec, tol, 50))
            eta = ctx.backward_error_dev(A, X, b)
            fwd = ((X - xs).abs().max() / xs.abs().max()).item()
            steps.append({"seed": s, "ms": ms, "method": method_name(rep),
                          "iterations": rep.iterations, "status": rep.status,
                          "backward_error": eta, "forward_error_vs_xstar": fwd,
                          "factor_ms": rep.factor_ms, "refine_ms": rep.refine_ms,
                          "perturbed_pivots": rep.perturbed_pivots})
        wall = time.perf_counter() - t_all0
        ms_ir = sum(st["ms"] for st in steps) / len(steps)
        all_ok = all(st["status"] == la_hip.LA_OK and st["backward_error"] <= tol for st in steps)
        value = flop * len(steps) / (sum(st["ms"] for st in steps) * 1e-3) / 1e12 if all_ok else None

        # ---------------- the other FP32 factor and plain FP64 on the same systems
        def per_system(prec, reps=1):
            res = []
            for s, A, b, xs in systems:
                ms, rep = timed(torch, lambda: ctx.gesv_ir_dev(A, b, X, prec, tol, 50))
                res.append({"seed": s, "ms": ms, "method": method_name(rep),
                            "iterations": rep.iterations, "status": rep.status,
                            "backward_error": ctx.backward_error_dev(A, X, b),
                            "forward_error_vs_xstar": ((X - xs).abs().max()
                                                       / xs.abs().max()).item()})
            return res
        others = {} if args.no_others else {
            m: per_system(la_hip.PREC_FP32 | o) for m, o in opt.items() if m != args.method}
        W = torch.empty((n, n), dtype=torch.float64, device="cuda")
        piv = torch.empty(n, dtype=torch.int64, device="cuda")
        info = torch.zeros(4, dtype=torch.int32, device="cuda")
        st64 = torch.zeros(1, dtype=torch.int32, device="cuda")
        X64 = torch.empty_like(X)
        fp64_runs = []
        for i, (s, A, b, xs) in enumerate(systems[:3]):
            W.copy_(A)
            if i == 0:       # warm-up of the FP64 kernels
                ctx.getrf_dev(W, piv, info)
                W.copy_(A)
            ms, _ = timed(torch, lambda: (ctx.getrf_dev(W, piv, info),
                                          ctx.getrs_dev(W, piv, b, X64, st64)))
            fp64_runs.append({"seed": s, "ms": ms, "backward_error":
                              ctx.backward_error_dev(A, X64, b),
                              "forward_error_vs_xstar": ((X64 - xs).abs().max()
                                                         / xs.abs().max()).item()})
        del W
        ms_64 = sum(r["ms"] for r in fp64_runs) / len(fp64_runs)

        # ---------------- instrumented solve: per-kernel-class HIP-event timing
        s0, A0, b0, _ = systems[0]
        ctx.reset_kernel_stats()
        ctx.set_profiling(True)
        ctx.gesv_ir_dev(A0, b0, X, head_prec, tol, 50)
        torch.cuda.synchronize()
        ctx.set_profiling(False)
        stats = ctx.kernel_stats()
        name, (calls, tot_ms, kflop, nbytes) = max(stats.items(), key=lambda kv: kv[1][1])
        avg_ms = tot_ms / calls
        if name.startswith("gemm"):
            peak = {"gemm64": PEAKS["fp64"], "gemm32": PEAKS["fp32"],
                    "gemm16": PEAKS["fp16"]}[name]
            achieved = kflop / calls / (avg_ms * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s"}
        else:
            achieved = nbytes / calls / (avg_ms * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": achieved, "peak": PEAKS["hbm_gbs"],
                    "unit": "GB/s"}
        roof["frac"] = roof["achieved"] / roof["peak"]
        key = f"{name}:{args.method}:n{n}"
        roof["traffic"] = load_traffic().get(key)
        roof.update({"kernel": name, "avg_launch_ms": avg_ms, "launches_per_step": calls,
                     "algorithmic_per_launch": kflop / calls if name.startswith("gemm")
                     else nbytes / calls})

        out = {
            "metric": METRIC, "value": value, "unit": "TFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_ir,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32-factor+fp64-refine",
            "data": "synthetic: seeded Householder-product A = U diag(sigma) V^T, normal b; "
                    f"steps cycle over seeds {used_seeds}",
            "config": {"workload": f"dense solve n={n}, kappa={kappa:g}, 1 RHS, FP32 LU "
                                   f"({ {'rbt': 'random butterfly, no pivoting', 'pivoted': 'partial pivoting', 'tournament': 'tournament pivoting'}[args.method]})"
                                   " + FP64 IR to eta<=1e-15*n",
                       "n": n, "kappa": kappa, "nrhs": 1, "seeds": used_seeds},
            "headline_steps": steps,
            "all_steps_converged": all_ok,
            "fp32_ir_other_factors": {
                m: {"mean_ms": sum(r["ms"] for r in runs) / len(runs),
                    "tflops": flop / (sum(r["ms"] for r in runs) / len(runs) * 1e-3) / 1e12,
                    "speedup_vs_fp64": ms_64 / (sum(r["ms"] for r in runs) / len(runs)),
                    "runs": runs} for m, runs in others.items()},
            "fp64_lu": {"mean_ms": ms_64, "tflops": flop / (ms_64 * 1e-3) / 1e12,
                        "frac_of_fp64_peak": flop / (ms_64 * 1e-3) / 1e12 / PEAKS["fp64"],
                        "runs": fp64_runs},
            "speedup_vs_fp64": ms_64 / ms_ir,
            "generate_ms": t_gen,
            "roofline": roof,
            "kernel_breakdown_ms": {k: round(v[1], 3) for k, v in stats.items()},
            "wall_s_timed": wall,
        }
        del systems, A0, b0

        # ---------------- kappa sweep over the family, every FP32 method
        if not args.no_sweep:
            A = torch.empty((n, n), dtype=torch.float64, device="cuda")
            b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
            xs = torch.empty_like(b)
            sweep = []
            for kap in (1e2, 1e4, 1e6):
                for s in SEEDS:
                    ctx.gen_svd_matrix_dev(n, kap, 64, s, A, b=b, xstar=xs)
                    row = {"kappa": kap, "seed": s}
                    for label, prec in (("pivoted", la_hip.PREC_FP32),
                                        ("butterfly", la_hip.PREC_FP32 | la_hip.IR_RBT),
                                        ("tournament", la_hip.PREC_FP32 | la_hip.IR_TOURNAMENT)):
                        ms, rep = timed(torch, lambda: ctx.gesv_ir_dev(A, b, X, prec, tol, 50))
                        row[label] = {"ms": round(ms, 2), "method": method_name(rep),
                                      "steps": rep.iterations, "status": rep.status,
                                      "eta": rep.backward_error}
                    sweep.append(row)
            labs = ("pivoted", "butterfly", "tournament")
            out["kappa_sweep"] = {
                "n": n, "runs": sweep,
                "served_by_own_factor": {lab: f"{sum(r[lab]['method'] == lab for r in sweep)}/{len(sweep)}"
                                         for lab in labs},
                "converged": {lab: f"{sum(r[lab]['status'] == 0 for r in sweep)}/{len(sweep)}"
                              for lab in labs},
                "median_ms": {lab: sorted(r[lab]["ms"] for r in sweep)[len(sweep) // 2]
                              for lab in labs}}
            # BF16 factor (per-column 2^k scaling) with GMRES refinement, kappa = 1e4
            ctx.gen_svd_matrix_dev(n, 1e4, 64, SEEDS[0], A, b=b, xstar=xs)
            ms, rep = timed(torch, lambda: ctx.gesv_ir_dev(
                A, b, X, la_hip.PREC_BF16 | la_hip.IR_GMRES, tol, 20))
            out["bf16_gmres_ir"] = {"kappa": 1e4, "ms": ms, "status": rep.status,
                                    "outer_steps": rep.iterations,
                                    "gmres_iterations": rep.inner_iterations,
                                    "backward_error": ctx.backward_error_dev(A, X, b),
                                    "forward_error_vs_xstar": ((X - xs).abs().max()
                                                               / xs.abs().max()).item(),
                                    "tflops": flop / (ms * 1e-3) / 1e12}
            del A
        if not args.no_cfg2:
            out["cfg2_n8192"] = cfg2(ctx, torch, 2)
            out["qr"] = qr_bench(ctx, torch)
            out["next_rows"] = next_rows(ctx, torch, 2)
        if not args.no_batched:
            out["batched_cholesky"] = batched_cholesky(ctx, args.steps, args.warmup)
            out["batched_lu_ir"] = batched_lu_ir(ctx, args.steps, args.warmup)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
        out["cpu_sweep"] = cpu_sweep()
        out["cpu_baseline"]["host_nproc"] = os.cpu_count()
        try:
            model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                     if l.startswith("model name")][0]
            out["cpu_baseline"]["host_cpu"] = model
        except Exception:
            pass
    print(json.dumps(out))


if __name__ == "__main__":
    main()
