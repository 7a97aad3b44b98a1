"""Batched uneven LU + FP64 iterative refinement (one persistent, graph-
replayed launch) against the oracle and the closed-form solutions.

Each problem is LUDecomposition::new + solve (src/decomp/lu.rs:104-278)
with refinement on top; the batch itself has no reference counterpart, so
the pins are: the generator against the oracle's Householder construction,
the plain-FP64 batch against the oracle's LU solve, and every refined
solution against its exact solution x* = V diag(1/sigma) U^T b and the
backward-error bound."""
import numpy as np
import pytest

import la_hip
import oracle

pytestmark = pytest.mark.gpu

SEED, BASE = 77, 5000


def make_batch(torch, ctx, sizes, nrhs, kappa=1e3, nrefl=64):
    sizes = np.asarray(sizes, dtype=np.int32)
    nrhs = np.asarray(nrhs, dtype=np.int32)
    n64 = sizes.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 ** 2)[:-1]]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]]).astype(np.int64)
    d = {k: torch.from_numpy(v).cuda() for k, v in
         dict(n=sizes, r=nrhs, oa=offA, ob=offB).items()}
    A = torch.empty(int((n64 ** 2).sum()), dtype=torch.float64, device="cuda")
    B = torch.empty(int((n64 * nrhs).sum()), dtype=torch.float64, device="cuda")
    Xs = torch.empty_like(B)
    ctx.gen_svd_batch_dev(d["n"], d["r"], d["oa"], d["ob"], kappa, nrefl, SEED, BASE, A, B, Xs)
    return dict(sizes=sizes, nrhs=nrhs, offA=offA, offB=offB, d=d, A=A, B=B, Xs=Xs)


def solve_batch(torch, ctx, bt, prec, tol=1e-14, max_iter=30, replays=1):
    plan = ctx.batch_plan(1, bt["sizes"], bt["nrhs"], prec)
    cnt = len(bt["sizes"])
    X = torch.full_like(bt["B"], float("nan"))
    st = torch.full((cnt,), -1, dtype=torch.int32, device="cuda")
    it = torch.full((cnt,), -1, dtype=torch.int32, device="cuda")
    be = torch.full((cnt,), -1.0, dtype=torch.float64, device="cuda")
    outs = []
    for _ in range(replays):
        ctx.gesv_ir_batched_dev(plan, bt["d"]["oa"], bt["d"]["ob"], bt["A"], bt["B"], X, tol,
                                max_iter, st, it, be)
        torch.cuda.synchronize()
        outs.append((X.cpu().numpy().copy(), st.cpu().numpy().copy(), it.cpu().numpy().copy(),
                     be.cpu().numpy().copy()))
    plan.close()
    return outs


def host_problem(bt, a, b, i):
    n, r = int(bt["sizes"][i]), int(bt["nrhs"][i])
    ai = a[bt["offA"][i]:bt["offA"][i] + n * n].reshape(n, n)
    bi = b[bt["offB"][i]:bt["offB"][i] + n * r].reshape(n, r)
    return ai, bi


def eta_of(ai, xi, bi):
    r = bi - ai @ xi
    return (np.abs(r).max(0) / (np.abs(ai).sum(1).max() * np.abs(xi).max(0)
                                + np.abs(bi).max(0))).max()


def test_generator_matches_oracle(tctx, torch_cuda):
    """A_i, B_i and x*_i from the batched generator equal the oracle's
    Householder construction with the documented per-problem streams."""
    torch = torch_cuda
    sizes, nrhs = [1, 5, 33, 64, 130], [1, 3, 8, 2, 1]
    bt = make_batch(torch, tctx, sizes, nrhs, kappa=1e3, nrefl=12)
    a, b, xs = bt["A"].cpu().numpy(), bt["B"].cpu().numpy(), bt["Xs"].cpu().numpy()
    for p, (n, r) in enumerate(zip(sizes, nrhs)):
        base = BASE + 256 * p
        ref, us, vs, sig = oracle.householder_svd_matrix(n, 1e3, 12, SEED, u_stream=base,
                                                         v_stream=base + 128)
        ai, bi = host_problem(bt, a, b, p)
        assert np.abs(ai - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), p
        bref = oracle.normal(SEED, base + 255, (n, r))
        # device log/sin/cos may differ from the host's libm in the last ulp
        assert np.abs(bi - bref).max() <= 1e-14 * np.abs(bref).max(), p
        xref = oracle.householder_exact_solution(us, vs, sig, bref)
        xi = xs[bt["offB"][p]:bt["offB"][p] + n * r].reshape(n, r)
        assert np.abs(xi - xref).max() <= 1e-12 * np.abs(xref).max(), p
        # the closed form: A x* = b
        assert np.abs(ai @ xi - bi).max() <= 1e-12 * np.abs(bi).max(), p


def test_fp64_batch_matches_oracle_lu(tctx, torch_cuda):
    """Plain FP64 batch: every solution agrees with the oracle's LU solve
    (lu.rs:104-278) to 1e-12 normwise, for uneven sizes across the panel
    width and the 64-row solve blocks."""
    torch = torch_cuda
    sizes = [1, 2, 7, 16, 17, 31, 32, 33, 64, 65, 100, 127, 128, 129, 200]
    nrhs = [(i % 8) + 1 for i in range(len(sizes))]
    bt = make_batch(torch, tctx, sizes, nrhs, kappa=1e2, nrefl=16)
    (x, st, it, be), = solve_batch(torch, tctx, bt, la_hip.PREC_FP64)
    assert (st == 0).all(), st
    a, b = bt["A"].cpu().numpy(), bt["B"].cpu().numpy()
    for i, (n, r) in enumerate(zip(sizes, nrhs)):
        ai, bi = host_problem(bt, a, b, i)
        lu, piv, _ = oracle.lu_new(ai)
        xref = oracle.lu_solve(lu, piv, bi)
        xi = x[bt["offB"][i]:bt["offB"][i] + n * r].reshape(n, r)
        err = np.abs(xi - xref).max() / np.abs(xref).max()
        assert err <= 1e-12, (i, n, err)
        # a residual at the rounding floor depends on the summation order:
        # compare the reported eta only down to that floor
        assert be[i] == pytest.approx(eta_of(ai, xi, bi), rel=0.5, abs=5e-17)


@pytest.mark.parametrize("prec", [la_hip.PREC_FP32, la_hip.PREC_FP16])
def test_lowprec_batch_refines_to_fp64(tctx, torch_cuda, prec):
    """FP32 / FP16 factorisations refined in FP64 reach eta <= 1e-14 for
    every problem (kappa = 1e3), the reported eta matches a host
    recomputation, the forward error is within kappa * eta of x*, and
    graph replays reproduce the first launch bit for bit."""
    torch = torch_cuda
    rng = np.random.default_rng(5)
    sizes = np.round(32 * 32 ** rng.random(48)).astype(np.int32)
    sizes[:4] = [1024, 1000, 32, 33]
    nrhs = rng.integers(1, 9, size=len(sizes)).astype(np.int32)
    bt = make_batch(torch, tctx, sizes, nrhs)
    outs = solve_batch(torch, tctx, bt, prec, replays=3)
    x, st, it, be = outs[0]
    assert (st == 0).all(), (st, be)
    assert (be <= 1e-14).all()
    assert (it >= 1).all() and (it <= 30).all()
    for o in outs[1:]:
        assert np.array_equal(o[0], x) and np.array_equal(o[1], st) and np.array_equal(o[3], be)
    a, b, xs = bt["A"].cpu().numpy(), bt["B"].cpu().numpy(), bt["Xs"].cpu().numpy()
    for i in range(len(sizes)):
        n, r = int(sizes[i]), int(nrhs[i])
        ai, bi = host_problem(bt, a, b, i)
        xi = x[bt["offB"][i]:bt["offB"][i] + n * r].reshape(n, r)
        xsi = xs[bt["offB"][i]:bt["offB"][i] + n * r].reshape(n, r)
        eta = eta_of(ai, xi, bi)
        assert eta <= 2e-14, (i, n, eta)
        assert be[i] == pytest.approx(eta, rel=0.5, abs=5e-16)
        fwd = np.abs(xi - xsi).max() / np.abs(xsi).max()
        assert fwd <= 1e3 * 1e3 * max(eta, 1e-16), (i, n, fwd)


def test_singular_and_bad_sizes(tctx, torch_cuda):
    """An exactly singular problem reports LA_SINGULAR without touching its
    neighbours; plan creation rejects orders above 1024 and nrhs above 8."""
    torch = torch_cuda
    sizes, nrhs = [40, 50, 60], [1, 2, 1]
    bt = make_batch(torch, tctx, sizes, nrhs)
    n1 = 50
    o = int(bt["offA"][1])
    A = bt["A"].view(-1)
    A[o:o + n1 * n1].view(n1, n1)[:, 7] = 0.0        # a zero column: U has a zero pivot
    (x, st, it, be), = solve_batch(torch, tctx, bt, la_hip.PREC_FP16)
    assert st[1] == la_hip.LA_SINGULAR
    assert st[0] == 0 and st[2] == 0
    with pytest.raises(Exception):
        tctx.batch_plan(1, np.array([1025], np.int32), np.array([1], np.int32), la_hip.PREC_FP16)
    with pytest.raises(Exception):
        tctx.batch_plan(1, np.array([10], np.int32), np.array([9], np.int32), la_hip.PREC_FP16)
    with pytest.raises(Exception):
        tctx.batch_plan(1, np.array([10], np.int32), np.array([1], np.int32), la_hip.PREC_BF16)


def test_full_batch_fp16_vs_fp64(tctx, torch_cuda):
    """The full workload: 4096 problems, n log-uniform in [32, 1024], kappa
    = 1e3, 1-8 right-hand sides; FP16 factorisation refined to eta <=
    1e-14 for every problem, compared with the batched plain FP64 LU."""
    torch = torch_cuda
    rng = np.random.default_rng(4096)
    sizes = np.round(32 * 32 ** rng.random(4096)).astype(np.int32)
    nrhs = rng.integers(1, 9, size=4096).astype(np.int32)
    bt = make_batch(torch, tctx, sizes, nrhs)
    (x16, st16, it16, be16), = solve_batch(torch, tctx, bt, la_hip.PREC_FP16)
    (x64, st64, it64, be64), = solve_batch(torch, tctx, bt, la_hip.PREC_FP64)
    assert (st16 == 0).all() and (st64 == 0).all()
    assert be16.max() <= 1e-14
    xs = bt["Xs"].cpu().numpy()
    # forward errors against x*, per problem, FP16-IR vs plain FP64
    for i in range(0, 4096, 61):
        n, r = int(sizes[i]), int(nrhs[i])
        sl = slice(bt["offB"][i], bt["offB"][i] + n * r)
        e16 = np.abs(x16[sl] - xs[sl]).max() / np.abs(xs[sl]).max()
        e64 = np.abs(x64[sl] - xs[sl]).max() / np.abs(xs[sl]).max()
        assert e16 <= 1e-10 and e64 <= 1e-10, (i, n, e16, e64)


def test_autotune_lu_grid(tctx, torch_cuda, tmp_path):
    """la_autotune kind 1: the grid (workgroups per CU) chosen per working
    precision is written to the JSON cache, loaded back by another context,
    reported by new plans, and a batch solved with the tuned grid is the same
    bit for bit as with the default grid (the grid only changes which
    workgroup takes which problem)."""
    import json
    torch = torch_cuda
    rng = np.random.default_rng(8)
    sizes = np.round(32 * 8 ** rng.random(40)).astype(np.int32)
    nrhs = rng.integers(1, 9, size=len(sizes)).astype(np.int32)
    bt = make_batch(torch, tctx, sizes, nrhs, nrefl=16)
    before = solve_batch(torch, tctx, bt, la_hip.PREC_FP32)[0]
    path = str(tmp_path / "lu_tune.json")
    tctx.autotune(1, 32, 256, path)
    with open(path) as f:
        cache = json.load(f)
    assert cache["kind"] == "lu-ir" and len(cache["workgroups_per_cu"]) == 3
    assert all(g >= 1 for g in cache["workgroups_per_cu"])
    after = solve_batch(torch, tctx, bt, la_hip.PREC_FP32)[0]
    for u, v in zip(before, after):
        assert np.array_equal(u, v)
    c2 = la_hip.Context(0)
    try:
        c2.autotune_load(path)
        plan = c2.batch_plan(1, sizes, nrhs, la_hip.PREC_FP16)
        assert "autotuned" in plan.describe()
        plan.close()
    finally:
        c2.close()
