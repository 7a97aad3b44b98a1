"""GPU tests of GMRES-based refinement (LA_IR_GMRES): corrections solve
A d = r by GMRES preconditioned with the low-precision factor, in FP64.

Pinned by closed forms like the other refinement tests: A = U diag(sigma)
V^T (Householder reflectors) has the exact solution x* = V diag(1/sigma)
U^T b, so the backward error reached (reported and recomputed) and the
forward error against x* are checked; the forward error must stay within
10 kappa eta.  BF16 factors at kappa = 1e4 and 1e6 are the cases plain
refinement cannot solve (kappa * u_bf16 >> 1, test_gpu_ir.py)."""
import pytest

import la_hip

pytestmark = pytest.mark.gpu


def _system(tctx, torch, n, kappa, seed, nrefl=64, nrhs=1):
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, nrhs), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, nrefl, seed, A, b=b, xstar=xs)
    return A, b, xs


def _check(tctx, A, b, X, xs, rep, tol, kappa):
    assert rep.status == la_hip.LA_OK, (rep.status, rep.backward_error, rep.iterations)
    assert rep.backward_error <= tol
    eta = tctx.backward_error_dev(A, X, b)
    assert eta <= tol, eta
    fwd = ((X - xs).abs().max() / xs.abs().max()).item()
    assert fwd <= 10 * kappa * max(eta, 1.1e-16), (fwd, eta)
    return eta, fwd


@pytest.mark.parametrize("kappa", [1e2, 1e4])
def test_bf16_gmres_ir_converges(tctx, torch_cuda, kappa):
    torch = torch_cuda
    n = 2000
    print(f"bf16 GMRES-IR n={n} kappa={kappa:g}", flush=True)
    A, b, xs = _system(tctx, torch, n, kappa, 31)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_BF16 | la_hip.IR_GMRES, tol, 20)
    print(f"  steps={rep.iterations} gmres={rep.inner_iterations} eta={rep.backward_error:.2e}",
          flush=True)
    # the first solve is BF16-accurate, so the low-precision factor ran
    assert rep.initial_backward_error > 1e-6
    assert rep.inner_iterations > 0
    _check(tctx, A, b, X, xs, rep, tol, kappa)


def test_bf16_gmres_ir_limit_kappa_1e6(tctx, torch_cuda):
    """kappa = 1e6 with a BF16 factor (kappa * u_bf16 ~ 4e3): the factor's
    error leaves the directions of all singular values below ~u_bf16
    ||A|| (about 60% of this geometric spectrum) unresolved, so GMRES needs
    a Krylov space of a sizeable fraction of n.  Within 8 corrections of at
    most 200 GMRES iterations each the solver must say that it did not
    converge, not return an inaccurate x as converged, and the backward
    error must still have fallen below the first solve's."""
    torch = torch_cuda
    n = 2000
    A, b, xs = _system(tctx, torch, n, 1e6, 31)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_BF16 | la_hip.IR_GMRES, tol, 8)
    assert rep.status == la_hip.LA_NOT_CONVERGED
    assert not (rep.backward_error <= tol)
    assert rep.backward_error < rep.initial_backward_error


def test_fp16_gmres_ir_converges(tctx, torch_cuda):
    torch = torch_cuda
    n, kappa = 2000, 1e5
    A, b, xs = _system(tctx, torch, n, kappa, 32)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP16 | la_hip.IR_GMRES, tol, 20)
    assert rep.inner_iterations > 0
    _check(tctx, A, b, X, xs, rep, tol, kappa)


@pytest.mark.parametrize("n", [1000, 2048])
def test_rbt_gmres_ir_serves_where_plain_refinement_falls_back(tctx, torch_cuda, n):
    """kappa = 1e6 systems on which refinement with the unpivoted butterfly
    factor stagnates (and falls back to partial pivoting); with GMRES
    corrections the butterfly factor itself serves."""
    torch = torch_cuda
    kappa = 1e6
    A, b, xs = _system(tctx, torch, n, kappa, 2025)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    plain = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_RBT, tol, 50)
    assert plain.status == la_hip.LA_OK
    assert plain.method == la_hip.IR_METHOD_RBT_THEN_PIVOTED
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_RBT | la_hip.IR_GMRES, tol, 50)
    print(f"  n={n} rbt+gmres method={rep.method} steps={rep.iterations} "
          f"gmres={rep.inner_iterations}", flush=True)
    assert rep.method == la_hip.IR_METHOD_RBT
    _check(tctx, A, b, X, xs, rep, tol, kappa)


def test_gmres_ir_multi_rhs(tctx, torch_cuda):
    torch = torch_cuda
    n, kappa = 1200, 1e6
    A, b, xs = _system(tctx, torch, n, kappa, 33, nrefl=16, nrhs=3)
    b[:, 1] *= 1e6
    xs[:, 1] *= 1e6
    X = torch.empty_like(b)
    tol = 1e-15 * n
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_BF16 | la_hip.IR_GMRES, tol, 20)
    assert rep.status == la_hip.LA_OK
    for c in range(3):
        eta = tctx.backward_error_dev(A, X[:, c:c + 1].contiguous(), b[:, c:c + 1].contiguous())
        assert eta <= tol, (c, eta)
        fwd = ((X[:, c] - xs[:, c]).abs().max() / xs[:, c].abs().max()).item()
        assert fwd <= 10 * kappa * max(eta, 1.1e-16), (c, fwd)
