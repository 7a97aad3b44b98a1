"""GPU tests of the Householder QR (src/decomp/qr.rs:26-237) and of
Matrix::pinverse (src/matrix/mod.rs:1049-1057) through the C ABI.

The GPU factor is the reference's (same sign rule, same Householder vectors,
same rdiag) computed blocked (32-column panels, the panel's reflections
applied as I - V T V^T on MFMA), so it agrees with the oracle to rounding:
the packed factor and rdiag within 1e-12 normwise relative to |A| (a
backward-stable factor of a well-conditioned random matrix), Q within
1e-12 of the oracle's Q and orthonormal to 1e-13 * max(m, n).  The solve is
the least-squares solve (the reference's literal formula is a defect,
pinned on the CPU in test_oracle.py): within 1e-10 of the oracle's corrected
solve and of LAPACK's least-squares solution, with the reference's status
order (b rows, then the full-rank scan)."""
import numpy as np
import pytest

import oracle
import la_hip
from refkit import m, approx_eq

pytestmark = pytest.mark.gpu

QR_KATS = [
    m([12.0, -51.0, 4.0], [6.0, 167.0, -68.0], [-4.0, 24.0, -41.0]),   # qr.rs:241-245
    m([1.0, 2.0], [3.0, 4.0], [5.0, 6.0]),                               # qr.rs:248-252
    m([1.0, 2.0, 3.0], [4.0, 5.0, 6.0]),                                 # qr.rs:255-259
]


def _nrel(x, ref, scale):
    return np.linalg.norm(x - ref) / max(scale, 1e-300)


@pytest.mark.parametrize("k", range(len(QR_KATS)))
def test_qr_kats(ctx, k):
    """qr.rs:241-259: (get_q() * get_r()).approx_eq(a), and the oracle's factor."""
    a = QR_KATS[k]
    qr, rd = ctx.qr_new(a)
    q = ctx.qr_q(qr, rd)
    r = oracle.qr_r(qr, rd)
    assert approx_eq(ctx.mul(q, r), a)
    qr_o, rd_o = oracle.qr_new(a)
    assert np.abs(qr - qr_o).max() <= 1e-12 * np.abs(a).max()
    assert np.abs(rd - rd_o).max() <= 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("shape", [(64, 64), (100, 60), (60, 100), (33, 33), (300, 300),
                                   (513, 257), (257, 513), (1000, 1000), (2048, 384)])
def test_qr_vs_oracle(ctx, shape):
    mm, nn = shape
    print(f"QR {shape}", flush=True)
    a = oracle.normal(4242, mm * 7 + nn, shape)
    qr, rd = ctx.qr_new(a)
    qr_o, rd_o = oracle.qr_new(a)
    scale = np.linalg.norm(a)
    assert np.array_equal(np.sign(rd), np.sign(rd_o)), "sign rule differs"
    assert _nrel(rd, rd_o, scale) <= 1e-12
    assert _nrel(qr, qr_o, scale) <= 1e-12
    q = ctx.qr_q(qr, rd)
    q_o = oracle.qr_q(qr_o, rd_o)
    k = min(mm, nn)
    assert np.linalg.norm(q - q_o) <= 1e-12 * np.sqrt(k)
    assert np.abs(q[:, :k].T @ q[:, :k] - np.eye(k)).max() <= 1e-13 * max(shape)
    assert np.all(q[:, k:] == 0.0)
    r = oracle.qr_r(qr, rd)
    assert np.linalg.norm(q @ r - a) <= 1e-13 * scale * np.sqrt(max(shape))


def test_qr_zero_column(ctx):
    """a = 0 for a column that is already zero below the diagonal: skipped
    (qr.rs:62), rdiag 0, the factor not of full rank, the solve None."""
    a = oracle.normal(5, 1, (70, 40))
    a[:, 35] = 0.0
    qr, rd = ctx.qr_new(a)
    qr_o, rd_o = oracle.qr_new(a)
    assert rd[35] == 0.0 and rd_o[35] == 0.0
    assert np.abs(qr - qr_o).max() <= 1e-12 * np.abs(a).max()
    x, st = ctx.qr_solve(qr, rd, oracle.normal(5, 2, (70, 1)))
    assert x is None and st == la_hip.LA_SINGULAR


@pytest.mark.parametrize("shape,nx", [((3, 3), 1), ((200, 200), 4), ((1000, 1000), 1),
                                      ((700, 300), 3), ((640, 640), 40)])
def test_qr_solve(ctx, shape, nx):
    mm, nn = shape
    a = oracle.normal(77, mm + nn, shape)
    b = oracle.normal(78, nx, (mm, nx))
    qr_o, rd_o = oracle.qr_new(a)
    x_o = oracle.qr_solve(qr_o, rd_o, b)
    qr, rd = ctx.qr_new(a)
    x, st = ctx.qr_solve(qr, rd, b)
    assert st == la_hip.LA_OK
    cond = np.linalg.cond(a)
    assert np.abs(x - x_o).max() <= 1e-13 * cond * max(np.abs(x_o).max(), 1.0) * np.sqrt(nn)
    x_ls = np.linalg.lstsq(a, b, rcond=None)[0]
    assert np.abs(x - x_ls).max() <= 1e-13 * cond * max(np.abs(x_ls).max(), 1.0) * np.sqrt(nn)


def test_qr_solve_status_order(ctx):
    """qr.rs:200 (b rows) before qr.rs:201-203 (full rank); is_full_rank
    reads past rdiag when n > m (a panic: LA_ERR_SHAPE)."""
    a = oracle.normal(1, 1, (5, 5))
    qr, rd = ctx.qr_new(a)
    x, st = ctx.qr_solve(qr, rd, np.ones((4, 1)))
    assert st == la_hip.LA_ERR_SHAPE
    rd0 = rd.copy()
    rd0[2] = 0.0
    x, st = ctx.qr_solve(qr, rd0, np.ones((5, 1)))
    assert st == la_hip.LA_SINGULAR
    qw, rw = ctx.qr_new(oracle.normal(1, 2, (3, 5)))
    x, st = ctx.qr_solve(qw, rw, np.ones((3, 1)))
    assert st == la_hip.LA_ERR_SHAPE


def test_pinverse(ctx):
    """mod.rs:1549-1552 (a.pinverse() * a ~ id(2)) and the oracle's pinverse."""
    a = m([1.0, 2.0], [3.0, 4.0], [5.0, 6.0])
    p, st = ctx.pinverse(a)
    assert st == la_hip.LA_OK
    assert approx_eq(ctx.mul(p, a), np.eye(2))
    for shape in [(200, 120), (512, 512), (1000, 257)]:
        b = oracle.normal(3, shape[0], shape)
        p, st = ctx.pinverse(b)
        p_o = oracle.pinverse(b)
        cond = np.linalg.cond(b)
        assert st == la_hip.LA_OK
        assert np.linalg.norm(p - p_o) <= 1e-13 * cond ** 2 * np.linalg.norm(p_o) * np.sqrt(shape[1])
        assert np.allclose(p, np.linalg.pinv(b), rtol=0, atol=1e-12 * cond ** 2 * np.abs(p_o).max())


def test_qr_device_entry_points(tctx, torch_cuda):
    """The _dev entry points (in place, device rdiag, Q and the solve) give
    the _rm results."""
    torch = torch_cuda
    a = oracle.normal(9, 3, (400, 300))
    b = oracle.normal(9, 4, (400, 2))
    qr, rd = tctx.qr_new(a)
    A = torch.from_numpy(a).cuda()
    RD = torch.empty(300, dtype=torch.float64, device="cuda")
    tctx.geqrf_dev(A, RD)
    assert np.abs(A.cpu().numpy() - qr).max() <= 1e-13 * np.abs(a).max()
    assert np.abs(RD.cpu().numpy() - rd).max() <= 1e-13 * np.abs(a).max()
    Q = torch.empty((400, 400), dtype=torch.float64, device="cuda")
    tctx.qr_q_dev(A, RD, Q)
    assert np.abs(Q.cpu().numpy() - tctx.qr_q(qr, rd)).max() <= 1e-13
    B = torch.from_numpy(b).cuda()
    X = torch.empty((300, 2), dtype=torch.float64, device="cuda")
    assert tctx.qrs_dev(A, RD, B, X) == la_hip.LA_OK
    x, _ = tctx.qr_solve(qr, rd, b)
    assert np.abs(X.cpu().numpy() - x).max() <= 1e-12 * np.abs(x).max()
    assert np.array_equal(B.cpu().numpy(), b)
    W = torch.from_numpy(oracle.normal(9, 5, (3, 5))).cuda()
    RW = torch.empty(3, dtype=torch.float64, device="cuda")
    tctx.geqrf_dev(W, RW)
    Xw = torch.empty((5, 1), dtype=torch.float64, device="cuda")
    assert tctx.qrs_dev(W, RW, torch.ones((3, 1), dtype=torch.float64, device="cuda"),
                        Xw) == la_hip.LA_ERR_SHAPE


@pytest.mark.parametrize("shape", [(300, 300), (513, 257), (2048, 384)])
def test_qr_column_kernel_path(ctx, shape, monkeypatch):
    """The per-column launches, which factor panels of more than 65536 rows
    (LAHIP_QR_COLUMN_KERNELS=1 selects them for every panel), against the
    oracle to the same bar as the cooperative panel."""
    monkeypatch.setenv("LAHIP_QR_COLUMN_KERNELS", "1")
    mm, nn = shape
    a = oracle.normal(4243, mm * 7 + nn, shape)
    qr, rd = ctx.qr_new(a)
    qr_o, rd_o = oracle.qr_new(a)
    scale = np.linalg.norm(a)
    assert np.array_equal(np.sign(rd), np.sign(rd_o))
    assert _nrel(rd, rd_o, scale) <= 1e-12
    assert _nrel(qr, qr_o, scale) <= 1e-12


def test_qr_panel_deterministic(ctx):
    """Every sum of the cooperative panel is taken in a fixed order: two runs
    give the same bits."""
    a = oracle.normal(4244, 1, (3000, 700))
    qr1, rd1 = ctx.qr_new(a)
    qr2, rd2 = ctx.qr_new(a)
    assert np.array_equal(qr1, qr2) and np.array_equal(rd1, rd2)
