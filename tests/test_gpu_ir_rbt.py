"""GPU tests of the butterfly form of the mixed-precision solver
(LA_IR_RBT): FP32 LU without pivoting of U^T [A 0; 0 I] V, refined in FP64
against A itself.

No reference counterpart (the FP64 solve it accelerates is
LUDecomposition::new + solve, src/decomp/lu.rs:104-278).  Pinned by the
closed form of the Householder-generated systems, x* = V diag(1/sigma) U^T b:
the normwise backward error must reach the tolerance, the forward error must
stay within kappa times it, and the butterfly attempt itself must converge
(a fallback to partial pivoting would hide a wrong unpivoted factor)."""
import numpy as np
import pytest

import la_hip

pytestmark = pytest.mark.gpu


def _system(torch, tctx, n, nrhs, kappa, seed):
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, nrhs), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, 16, seed, A, b=b, xstar=xs)
    return A, b, xs


@pytest.mark.parametrize("n,kappa", [(1000, 1e2), (1001, 1e4), (1002, 1e6), (1536, 1e6)])
def test_rbt_converges_without_fallback(tctx, torch_cuda, n, kappa):
    torch = torch_cuda
    A, b, xs = _system(torch, tctx, n, 1, kappa, 21)
    X = torch.full_like(b, float("nan"))
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_RBT, tol, 30)
    assert rep.status == la_hip.LA_OK, (rep.status, rep.backward_error, rep.method)
    assert rep.method == la_hip.IR_METHOD_RBT
    assert rep.iterations <= 10, rep.iterations
    eta = tctx.backward_error_dev(A, X, b)
    assert eta <= tol
    assert rep.backward_error == pytest.approx(eta, rel=0.5, abs=1e-17)
    fwd = ((X - xs).abs().max() / xs.abs().max()).item()
    assert fwd <= 10 * kappa * max(eta, 1.1e-16), fwd


@pytest.mark.parametrize("n", [1, 2, 3, 5, 63, 64, 65, 129, 255, 257])
def test_rbt_small_and_ragged_orders(tctx, torch_cuda, n):
    """Orders that are not multiples of 4 (identity padding) or of the
    64-wide blocks."""
    torch = torch_cuda
    A, b, xs = _system(torch, tctx, n, 2, 1e3, 3 + n)
    X = torch.empty_like(b)
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_RBT, 1e-15 * max(n, 16), 30)
    assert rep.status == la_hip.LA_OK and rep.method == la_hip.IR_METHOD_RBT, \
        (rep.status, rep.method, rep.backward_error)
    fwd = ((X - xs).abs().max() / xs.abs().max()).item()
    assert fwd <= 1e-9, fwd


@pytest.mark.parametrize("nrhs", [3, 16])
def test_rbt_multi_rhs_matches_pivoted(tctx, torch_cuda, nrhs):
    """Several right-hand sides: the butterfly and the pivoted solvers agree
    to within kappa times the tolerance they both reach."""
    torch = torch_cuda
    n = 777
    A, b, xs = _system(torch, tctx, n, nrhs, 1e3, 8)
    X1 = torch.empty_like(b)
    X2 = torch.empty_like(b)
    r1 = tctx.gesv_ir_dev(A, b, X1, la_hip.PREC_FP32 | la_hip.IR_RBT, 1e-16 * n, 30)
    r2 = tctx.gesv_ir_dev(A, b, X2, la_hip.PREC_FP32, 1e-16 * n, 30)
    assert r1.status == la_hip.LA_OK and r1.method == la_hip.IR_METHOD_RBT
    assert r2.status == la_hip.LA_OK and r2.method == la_hip.IR_METHOD_PIVOTED
    d = ((X1 - X2).abs().max() / X2.abs().max()).item()
    assert d <= 10 * 1e3 * 1e-16 * n, d


def test_rbt_singular_falls_back_and_reports(tctx, torch_cuda):
    """An exactly singular matrix: the butterfly attempt cannot converge, the
    pivoted attempt finds the zero pivot, and the report says both."""
    torch = torch_cuda
    n = 300
    A, b, xs = _system(torch, tctx, n, 1, 1e2, 4)
    A[:, 17] = 0.0
    X = torch.empty_like(b)
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_RBT, 1e-15 * n, 8)
    assert rep.method == la_hip.IR_METHOD_RBT_THEN_PIVOTED
    assert rep.status in (la_hip.LA_SINGULAR, la_hip.LA_NOT_CONVERGED)


@pytest.mark.parametrize("n,kappa", [(2000, 1e6), (1984, 1e4)])
def test_rbt_fallback_still_meets_the_tolerance(tctx, torch_cuda, n, kappa):
    """Systems on which the unpivoted attempt stagnates (measured): the
    attempt is abandoned early, the pivoted factor finishes the solve, the
    result meets the tolerance and the report names both attempts."""
    torch = torch_cuda
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, 64, 2025, A, b=b, xstar=xs)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_RBT, tol, 50)
    assert rep.status == la_hip.LA_OK
    assert rep.method == la_hip.IR_METHOD_RBT_THEN_PIVOTED
    assert tctx.backward_error_dev(A, X, b) <= tol
    fwd = ((X - xs).abs().max() / xs.abs().max()).item()
    assert fwd <= 10 * kappa * 1e-15 * n


def test_rbt_rejects_other_precisions(tctx, torch_cuda):
    torch = torch_cuda
    A, b, xs = _system(torch, tctx, 64, 1, 1e2, 4)
    X = torch.empty_like(b)
    with pytest.raises(Exception):
        tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP16 | la_hip.IR_RBT, 1e-13, 8)
