"""Batched uneven Cholesky solve on the GPU (one persistent, graph-replayed
launch) against the CPU oracle, problem by problem.

Each problem restates CholeskyDecomposition::new + solve
(src/decomp/cholesky.rs:56-144); failing problems (asymmetric, not positive
definite) must report their own status without disturbing neighbours."""
import numpy as np
import pytest

import oracle
import la_hip

pytestmark = pytest.mark.gpu




def _pack(mats, rhs):
    offA, offB, a_parts, b_parts = [], [], [], []
    oa = ob = 0
    for m, b in zip(mats, rhs):
        offA.append(oa); offB.append(ob)
        a_parts.append(m.ravel()); b_parts.append(b.ravel())
        oa += m.size; ob += b.size
    return (np.array(offA, np.int64), np.array(offB, np.int64),
            np.concatenate(a_parts), np.concatenate(b_parts))


def _run(tctx, torch, sizes, mats, rhs, replays=1):
    offA, offB, a, b = _pack(mats, rhs)
    d = lambda x: torch.from_numpy(x).cuda()
    dA, dB, dOA, dOB = d(a), d(b), d(offA), d(offB)
    dL = torch.full_like(dA, np.nan)
    dX = torch.full_like(dB, np.nan)
    st = torch.full((len(sizes),), -1, dtype=torch.int32, device="cuda")
    plan = tctx.batch_plan(0, sizes)
    outs = []
    for _ in range(replays):
        dX.fill_(np.nan)
        tctx.potrf_potrs_batched_dev(plan, dOA, dOB, dA, dL, dB, dX, st)
        torch.cuda.synchronize()
        outs.append((dL.cpu().numpy().copy(), dX.cpu().numpy().copy(), st.cpu().numpy().copy()))
    plan.close()
    return offA, offB, outs


@pytest.mark.parametrize("nb,rt", [(8, 1), (8, 2), (16, 1), (16, 2)])
def test_batched_cholesky_vs_oracle(tctx, torch_cuda, tmp_path, nb, rt):
    """Every problem against the oracle, with each tile configuration the
    autotuner can choose (panel width, row tiles per MFMA pass) forced for
    all size buckets."""
    torch = torch_cuda
    cache = tmp_path / "tune.json"
    cache.write_text('{"kind": "cholesky", "buckets": [32, 64, 128, 256], '
                     f'"panel_width": [{nb}, {nb}, {nb}, {nb}], '
                     f'"row_tiles": [{rt}, {rt}, {rt}, {rt}]}}\n')
    tctx.autotune_load(str(cache))
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in rng.integers(1, 257, size=120)] + [1, 2, 31, 32, 33, 64, 65, 128, 129, 256]
    mats = [oracle.spd_matrix(n, 77, 1000 + i) for i, n in enumerate(sizes)]
    rhs = [oracle.normal(77, 50000 + i, (n,)) for i, n in enumerate(sizes)]
    offA, offB, outs = _run(tctx, torch, sizes, mats, rhs, replays=3)
    for L, X, st in outs:                      # every graph replay is checked
        assert np.all(st == 0)
        for i, n in enumerate(sizes):
            l_o, so = oracle.chol_new(mats[i])
            x_o = oracle.chol_solve(l_o, rhs[i].reshape(-1, 1)).ravel()
            l = L[offA[i]:offA[i] + n * n].reshape(n, n)
            x = X[offB[i]:offB[i] + n]
            assert np.abs(l - l_o).max() <= 1e-12 * np.abs(l_o).max(), (i, n)
            assert np.all(np.triu(l, 1) == 0.0)
            assert np.abs(x - x_o).max() <= 1e-12 * max(np.abs(x_o).max(), 1e-300), (i, n)


def test_batched_failures_are_isolated(tctx, torch_cuda):
    torch = torch_cuda
    sizes = [40, 100, 200, 7, 150]
    mats = [oracle.spd_matrix(n, 3, 10 + i) for i, n in enumerate(sizes)]
    mats[1] = mats[1].copy(); mats[1][70, 3] = np.nextafter(mats[1][70, 3], 9.0)   # asymmetric
    mats[2] = mats[2].copy(); mats[2][150, 150] = -5.0                          # not SPD
    rhs = [oracle.normal(3, 90 + i, (n,)) for i, n in enumerate(sizes)]
    offA, offB, outs = _run(tctx, torch, sizes, mats, rhs)
    L, X, st = outs[0]
    assert st.tolist() == [0, la_hip.LA_NOT_SYMMETRIC, la_hip.LA_NOT_SPD, 0, 0]
    for i in (0, 3, 4):
        n = sizes[i]
        l_o, _ = oracle.chol_new(mats[i])
        x_o = oracle.chol_solve(l_o, rhs[i].reshape(-1, 1)).ravel()
        assert np.abs(X[offB[i]:offB[i] + n] - x_o).max() <= 1e-12 * np.abs(x_o).max()


def test_full_batch_20000_properties(tctx, torch_cuda):
    """The full workload: 20000 SPD problems, n ~ U{8..256},
    A_i = G G^T / n + I generated on the GPU; every solution's normwise
    backward error is checked on the host."""
    torch = torch_cuda
    rng = np.random.default_rng(20000)
    sizes = rng.integers(8, 257, size=20000).astype(np.int32)
    offA = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)[:-1]])
    offB = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))[:-1]])
    tot_a, tot_b = int((sizes.astype(np.int64) ** 2).sum()), int(sizes.sum())
    dn, dOA, dOB = (torch.from_numpy(x).cuda() for x in (sizes, offA, offB))
    A = torch.empty(tot_a, dtype=torch.float64, device="cuda")
    G = torch.empty_like(A)
    tctx.gen_spd_batch_dev(dn, dOA, 4, 1000, A, G)
    del G
    B = torch.empty(tot_b, dtype=torch.float64, device="cuda")
    tctx.fill_normal_dev(4, 7, B)
    L = torch.empty_like(A)
    X = torch.empty_like(B)
    st = torch.empty(len(sizes), dtype=torch.int32, device="cuda")
    plan = tctx.batch_plan(0, sizes)
    tctx.potrf_potrs_batched_dev(plan, dOA, dOB, A, L, B, X, st)
    torch.cuda.synchronize()
    assert int((st != 0).sum().item()) == 0
    a, b, x = A.cpu().numpy(), B.cpu().numpy(), X.cpu().numpy()
    worst = 0.0
    for i in range(0, len(sizes), 37):          # a spread sample of the problems
        n = int(sizes[i])
        ai = a[offA[i]:offA[i] + n * n].reshape(n, n)
        xi, bi = x[offB[i]:offB[i] + n], b[offB[i]:offB[i] + n]
        eta = np.abs(bi - ai @ xi).max() / (np.abs(ai).sum(1).max() * np.abs(xi).max() + np.abs(bi).max())
        worst = max(worst, eta)
    assert worst < 1e-15
    plan.close()
