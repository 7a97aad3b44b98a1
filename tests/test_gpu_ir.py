"""GPU tests of the mixed-precision solver and the seeded generator.

The generator is pinned against the oracle's numpy restatement; the FP32
factorisation with FP64 refinement is pinned by closed forms: for
A = U diag(sigma) V^T the exact solution is x* = V diag(1/sigma) U^T b,
and the normwise backward error must reach the requested tolerance."""
import numpy as np
import pytest

import oracle
import la_hip

pytestmark = pytest.mark.gpu




def test_generator_matches_oracle(tctx, torch_cuda):
    torch = torch_cuda
    n, kappa, nrefl, seed = 200, 1e4, 8, 7
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 2), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, nrefl, seed, A, b=b, xstar=xs)
    torch.cuda.synchronize()
    a_o, us, vs, sig = oracle.householder_svd_matrix(n, kappa, nrefl, seed)
    b_o = oracle.normal(seed, 300, (n, 2))
    x_o = oracle.householder_exact_solution(us, vs, sig, b_o)
    assert np.abs(A.cpu().numpy() - a_o).max() < 1e-13
    assert np.abs(b.cpu().numpy() - b_o).max() < 1e-13
    assert np.abs(xs.cpu().numpy() - x_o).max() / np.abs(x_o).max() < 1e-12


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_fp32_ir_converges(tctx, torch_cuda, kappa):
    torch = torch_cuda
    n = 1000
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, 16, 11, A, b=b, xstar=xs)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32, tol, 30)
    assert rep.status == la_hip.LA_OK, (rep.status, rep.backward_error)
    assert rep.backward_error <= tol
    # independent check of eta and of the forward error against x*
    eta = tctx.backward_error_dev(A, X, b)
    assert eta <= tol
    # forward error is bounded by kappa times the backward error reached
    fwd = ((X - xs).abs().max() / xs.abs().max()).item()
    assert fwd <= 10 * kappa * max(eta, 1.1e-16)
    assert rep.iterations <= 12
    # refined to the FP64 floor, the solution agrees with x* to ~kappa*u
    rep2 = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32, 4e-16, 40)
    fwd2 = ((X - xs).abs().max() / xs.abs().max()).item()
    assert fwd2 <= 50 * kappa * 1.1e-16, (fwd2, rep2.backward_error, rep2.iterations)


@pytest.mark.parametrize("nrhs", [1, 3, 16])
def test_fp32_ir_multi_rhs(tctx, torch_cuda, nrhs):
    torch = torch_cuda
    n = 777
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, nrhs), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, 1e3, 8, 5, A, b=b, xstar=xs)
    X = torch.empty_like(b)
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32, 1e-16 * n, 30)
    assert rep.status == la_hip.LA_OK
    eta = tctx.backward_error_dev(A, X, b)
    assert eta <= 1e-16 * n
    assert ((X - xs).abs().max() / xs.abs().max()).item() < 10 * 1e3 * eta


def test_fp64_full_size_cfg2(tctx, torch_cuda):
    """Config 2 at full size (n = 8192, 16 right-hand sides, normal entries):
    size-independent properties - pivot vector is a permutation, P A = L U
    on random probes, normwise backward error of the solve."""
    torch = torch_cuda
    n = 8192
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    tctx.fill_normal_dev(2024, 0, A)
    B = torch.empty((n, 16), dtype=torch.float64, device="cuda")
    tctx.fill_normal_dev(2024, 2, B)
    W = A.clone()
    piv = torch.empty(n, dtype=torch.int64, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    tctx.getrf_dev(W, piv, info)
    X = torch.empty_like(B)
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    tctx.getrs_dev(W, piv, B, X, st)
    torch.cuda.synchronize()
    assert info[1].item() == 0 and st.item() == 0
    p = piv.cpu()
    assert bool((p.sort().values == torch.arange(n)).all())
    z = torch.randn(n, 2, dtype=torch.float64, device="cuda")
    L = torch.tril(W, -1) + torch.eye(n, dtype=torch.float64, device="cuda")
    U = torch.triu(W)
    r = A[piv] @ z - L @ (U @ z)
    # classical bound ||PA - LU|| <= gamma_n ||L|| ||U||; n*u = 9.1e-13 here
    assert (r.abs().max() / (A.abs().sum(1).max() * z.abs().max())).item() < 2e-13
    eta = tctx.backward_error_dev(A, X, B)
    assert eta < 1e-14


@pytest.mark.parametrize("prec,kappa,floor_lo,floor_hi", [
    (la_hip.PREC_BF16, 1e2, 1e-5, 3e-2),
    (la_hip.PREC_FP16, 1e2, 1e-6, 3e-3),
    (la_hip.PREC_FP16, 1e3, 1e-6, 3e-3),
])
def test_lowprec_ir_converges(tctx, torch_cuda, prec, kappa, floor_lo, floor_hi):
    """BF16/FP16-operand trailing updates with per-column 2^k scaling: the
    first solve's backward error sits at the operand precision (so the
    low-precision MFMA path is the one that ran), and FP64 refinement reaches
    1e-15*n where kappa * u < 1."""
    torch = torch_cuda
    n = 1500
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, 16, 21, A, b=b, xstar=xs)
    X = torch.empty_like(b)
    rep = tctx.gesv_ir_dev(A, b, X, prec, 1e-15 * n, 200)
    assert floor_lo < rep.initial_backward_error < floor_hi, rep.initial_backward_error
    assert rep.status == la_hip.LA_OK, (rep.status, rep.backward_error, rep.iterations)
    eta = tctx.backward_error_dev(A, X, b)
    assert eta <= 1e-15 * n
    assert ((X - xs).abs().max() / xs.abs().max()).item() <= 10 * kappa * max(eta, 1.1e-16)


def test_bf16_ir_reports_non_convergence(tctx, torch_cuda):
    """kappa * u_bf16 >> 1: plain refinement cannot converge; the solver must
    say so instead of returning a silently inaccurate x."""
    torch = torch_cuda
    n = 1000
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, 1e6, 16, 22, A, b=b, xstar=xs)
    X = torch.empty_like(b)
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_BF16, 1e-15 * n, 10)
    assert rep.status in (la_hip.LA_NOT_CONVERGED, la_hip.LA_SINGULAR)
    assert not (rep.backward_error <= 1e-15 * n)


def _eta_per_column(a, x, b):
    """max over columns c of ||b_c - A x_c||inf / (||A||inf ||x_c||inf + ||b_c||inf)"""
    r = np.abs(b - a @ x).max(axis=0)
    an = np.abs(a).sum(axis=1).max()
    return float((r / (an * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))).max())


def test_ir_backward_error_is_per_column(tctx, torch_cuda):
    """Right-hand sides of very different scales: the reported backward error
    must be the maximum of the per-column backward errors (a norm taken over
    all columns together would hide a small-scale column that is still far
    from converged)."""
    torch = torch_cuda
    n, nrhs = 600, 3
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, nrhs), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, 1e4, 8, 17, A, b=b, xstar=xs)
    b[:, 1] *= 1e8
    b[:, 2] *= 1e-8
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32, tol, 30)
    torch.cuda.synchronize()
    a, x, bb = A.cpu().numpy(), X.cpu().numpy(), b.cpu().numpy()
    host = _eta_per_column(a, x, bb)
    assert rep.status == la_hip.LA_OK and rep.backward_error <= tol
    # a converged residual is rounding noise: device and host agree on the
    # bound, not digit for digit
    assert host <= tol, host
    assert tctx.backward_error_dev(A, X, b) <= tol
    # a perturbed small-scale column is seen even though the others are
    # converged: its residual dominates its own column's backward error
    X[:, 2] *= 1.0 + 1e-6
    torch.cuda.synchronize()
    host2 = _eta_per_column(a, X.cpu().numpy(), bb)
    dev2 = tctx.backward_error_dev(A, X, b)
    # eta of column 2 is 1e-6 relative perturbation / effective condition,
    # far above tol; a norm over all columns together would stay near tol
    assert host2 > 10 * tol and abs(dev2 - host2) <= 0.05 * host2, (host2, dev2)
