"""The allocation-free LU entry point (la_dgetrf_noalloc_dev) against
la_dgetrf_dev through the C ABI: identical packed factors, permutation
vector and sign, on two calls in a row (the second smaller, so the shared
scratch holds the first call's contents), and the size limit as an error
returned before any work."""
import pytest

import la_hip

pytestmark = pytest.mark.gpu


def _factor(tctx, torch, A, noalloc):
    m, n = A.shape
    W = A.clone()
    piv = torch.empty(m, dtype=torch.int64, device="cuda")
    info = torch.full((4,), -7, dtype=torch.int32, device="cuda")
    (tctx.getrf_noalloc_dev if noalloc else tctx.getrf_dev)(W, piv, info)
    torch.cuda.synchronize()
    return W, piv, info


@pytest.mark.parametrize("shape", [(3000, 3000), (700, 650)])
def test_noalloc_matches_getrf_dev(tctx, torch_cuda, shape):
    torch = torch_cuda
    print(f"noalloc getrf {shape}", flush=True)
    A = torch.empty(shape, dtype=torch.float64, device="cuda")
    tctx.fill_uniform_dev(97, shape[0], A)
    W0, p0, i0 = _factor(tctx, torch, A, noalloc=False)
    W1, p1, i1 = _factor(tctx, torch, A, noalloc=True)
    assert int(i1[1].item()) == la_hip.LA_OK
    assert torch.equal(p0, p1)
    assert torch.equal(i0[:2], i1[:2])
    assert torch.equal(W0, W1)


def test_noalloc_size_limit(tctx, torch_cuda):
    torch = torch_cuda
    A = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    piv = torch.zeros(1, dtype=torch.int64, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    lib = tctx.lib
    st = lib.la_dgetrf_noalloc_dev(tctx.h, 65537, 65537, la_hip._ptr(A), 65537,
                                   la_hip._ptr(piv), la_hip._ptr(info))
    assert st == la_hip.LA_ERR_ARG
