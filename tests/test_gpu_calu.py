"""GPU tests of the tournament-pivoted (CALU) FP32 factor of the refinement
solver (LA_IR_TOURNAMENT): pivots of every 64-column panel chosen by partial
pivoting along a reduction tree of candidate rows, the panel then factored
without pivoting.

Pinned by closed forms like the other refinement tests (A = U diag(sigma)
V^T has x* = V diag(1/sigma) U^T b): the tournament factor must be the one
that converged (a wrong factor would be hidden by the fallback to partial
pivoting otherwise), its first solve must be about as accurate as the
partial-pivoting factor's (growth of the same order), and refinement must
reach eta <= 1e-15 n with the forward error inside the perturbation bound."""
import pytest

import la_hip

pytestmark = pytest.mark.gpu


def _system(tctx, torch, n, kappa, seed, nrefl=64, nrhs=1):
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, nrhs), dtype=torch.float64, device="cuda")
    xs = torch.empty_like(b)
    tctx.gen_svd_matrix_dev(n, kappa, nrefl, seed, A, b=b, xstar=xs)
    return A, b, xs


@pytest.mark.parametrize("n,kappa", [(1000, 1e2), (1000, 1e6), (2048, 1e6), (3000, 1e4),
                                     (777, 1e6), (65, 1e3), (130, 1e5), (257, 1e6),
                                     (4097, 1e6)])
def test_tournament_factor_serves(tctx, torch_cuda, n, kappa):
    torch = torch_cuda
    print(f"tournament n={n} kappa={kappa:g}", flush=True)
    A, b, xs = _system(tctx, torch, n, kappa, 2025)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    ref = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32, tol, 50)
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_TOURNAMENT, tol, 50)
    print(f"  method={la_hip.IR_METHOD_NAMES[rep.method]} steps={rep.iterations} "
          f"eta0={rep.initial_backward_error:.2e} (pivoted {ref.initial_backward_error:.2e})",
          flush=True)
    assert rep.status == la_hip.LA_OK and rep.backward_error <= tol
    assert rep.method == la_hip.IR_METHOD_TOURNAMENT
    assert rep.initial_backward_error <= 20 * ref.initial_backward_error + 1e-16
    eta = tctx.backward_error_dev(A, X, b)
    assert eta <= tol
    fwd = ((X - xs).abs().max() / xs.abs().max()).item()
    kinf = n * kappa
    assert fwd <= 2 * kinf * eta / (1 - kinf * eta)


def test_tournament_multi_rhs_and_gmres(tctx, torch_cuda):
    torch = torch_cuda
    n, kappa = 1500, 1e6
    A, b, xs = _system(tctx, torch, n, kappa, 41, nrefl=16, nrhs=3)
    X = torch.empty_like(b)
    tol = 1e-15 * n
    for extra in (0, la_hip.IR_GMRES):
        rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_TOURNAMENT | extra, tol, 50)
        assert rep.status == la_hip.LA_OK
        assert rep.method == la_hip.IR_METHOD_TOURNAMENT
        assert tctx.backward_error_dev(A, X, b) <= tol


def test_tournament_rejects_other_precisions(tctx, torch_cuda):
    torch = torch_cuda
    A, b, xs = _system(tctx, torch, 100, 10.0, 1, nrefl=4)
    X = torch.empty_like(b)
    with pytest.raises(la_hip.LaError) as e:
        tctx.gesv_ir_dev(A, b, X, la_hip.PREC_BF16 | la_hip.IR_TOURNAMENT, 1e-13, 10)
    assert e.value.status == la_hip.LA_ERR_ARG


def test_tournament_singular_reports(tctx, torch_cuda):
    """An exactly singular A: neither factor can serve; the report says so."""
    torch = torch_cuda
    n = 300
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    tctx.fill_uniform_dev(5, 0, A)
    A[:, 7] = 0.0
    b = torch.ones((n, 1), dtype=torch.float64, device="cuda")
    X = torch.empty_like(b)
    rep = tctx.gesv_ir_dev(A, b, X, la_hip.PREC_FP32 | la_hip.IR_TOURNAMENT, 1e-15 * n, 20)
    assert rep.status != la_hip.LA_OK
    assert rep.method == la_hip.IR_METHOD_TOURNAMENT_THEN_PIVOTED
