"""The LU look-ahead schedule against the plain single-stream schedule.

With look-ahead the panels of outer block j+1 (src/decomp/lu.rs:104-168's
column loop, 64 columns at a time) run on a stream confined to a few CUs of
every XCD while the rest of the GPU applies block j to the columns right of
block j+1.  Every location still sees the same kernels with the same
operands in the same order, so the factor, the pivots, the pivot sign and
the refined solutions must be bit-identical to the plain schedule's.
LAHIP_LOOKAHEAD=0 selects the plain schedule; the library reads it on every
call.  The parity tests against the oracle run with look-ahead on (the
default) for every n >= 512."""
import os

import pytest

import la_hip

pytestmark = pytest.mark.gpu


def _set(lookahead):
    os.environ["LAHIP_LOOKAHEAD"] = "1" if lookahead else "0"


@pytest.fixture(autouse=True)
def _restore_env():
    yield
    os.environ.pop("LAHIP_LOOKAHEAD", None)


def _getrf(tctx, torch, A0, lookahead):
    _set(lookahead)
    A = A0.clone()
    piv = torch.empty(A0.shape[0], dtype=torch.int64, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    tctx.getrf_dev(A, piv, info)
    torch.cuda.synchronize()
    return A, piv, info


# (rows, cols): square sizes with two to 17 outer blocks, tall and wide
# shapes, and n = 8448 whose first block has more rows (8448) than the
# panel stream's CUs hold in FP64 (8192), so the plain schedule runs first
# and hands over to the look-ahead schedule.
@pytest.mark.parametrize("shape", [(512, 512), (600, 600), (1000, 1000), (2048, 2048),
                                   (4100, 4100), (1500, 900), (900, 1500), (8448, 8448)])
def test_lookahead_getrf_bit_identical(tctx, torch_cuda, shape):
    torch = torch_cuda
    print(f"look-ahead getrf {shape}", flush=True)
    A0 = torch.empty(shape, dtype=torch.float64, device="cuda")
    tctx.fill_normal_dev(4242, shape[0] + 3 * shape[1], A0)
    torch.cuda.synchronize()
    A1, p1, i1 = _getrf(tctx, torch, A0, False)
    A2, p2, i2 = _getrf(tctx, torch, A0, True)
    assert i1[1].item() == 0 and i2[1].item() == 0
    assert torch.equal(p1, p2)
    assert i1[0].item() == i2[0].item()
    assert torch.equal(A1, A2)


@pytest.mark.parametrize("prec,n,kappa", [(la_hip.PREC_FP32, 3000, 1e6),
                                          (la_hip.PREC_FP32, 4608, 1e4),
                                          (la_hip.PREC_BF16, 2500, 1e2)])
def test_lookahead_refinement_bit_identical(tctx, torch_cuda, prec, n, kappa):
    """The low-precision factors of la_dgesv_ir_dev (FP32 panel, FP32 or BF16
    trailing update) go through the same schedule."""
    torch = torch_cuda
    print(f"look-ahead IR prec {prec} n {n}", flush=True)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 2), dtype=torch.float64, device="cuda")
    xs = torch.empty((n, 2), dtype=torch.float64, device="cuda")
    tctx.gen_svd_matrix_dev(n, kappa, 16, 97, A, b=b, xstar=xs)
    out = []
    for la in (False, True):
        _set(la)
        X = torch.empty((n, 2), dtype=torch.float64, device="cuda")
        rep = tctx.gesv_ir_dev(A, b, X, prec, 1e-16 * n, 50)
        torch.cuda.synchronize()
        out.append((X, rep.iterations, rep.status, rep.backward_error))
    (X1, it1, st1, e1), (X2, it2, st2, e2) = out
    assert st1 == st2 == la_hip.LA_OK
    assert it1 == it2 and e1 == e2
    assert torch.equal(X1, X2)


@pytest.mark.parametrize("shape", [(3000, 3000), (8448, 8448)])
def test_lookahead_repeatable(tctx, torch_cuda, shape):
    """Four look-ahead factorisations of one matrix all equal the
    single-stream factor bit for bit.  The two streams share buffers only
    through events (swap maps and L11^-1 blocks alternate between outer
    blocks), so the result may not depend on how the streams' kernels
    happen to interleave from run to run."""
    torch = torch_cuda
    print(f"look-ahead repeat {shape}", flush=True)
    A0 = torch.empty(shape, dtype=torch.float64, device="cuda")
    tctx.fill_normal_dev(777, shape[0], A0)
    torch.cuda.synchronize()
    A1, p1, _ = _getrf(tctx, torch, A0, False)
    for _ in range(4):
        A2, p2, i2 = _getrf(tctx, torch, A0, True)
        assert i2[1].item() == 0
        assert torch.equal(p1, p2) and torch.equal(A1, A2)
