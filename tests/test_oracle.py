"""Pins the CPU oracle (oracle/la_oracle.c) before anything is compared to it.

Every known-answer test the reference holds for the hot path is replayed
here bit-for-bit (file:line of the reference test in each docstring), and
the parts those leave open are pinned by independent routes: LAPACK's
dgetrf through scipy (same partial-pivoting rule, different code), numpy's
Cholesky and matmul, closed forms on 2x2 inputs, and invariants
(P*A = L*U, permutation parity, L*L^T = A)."""
import numpy as np
import pytest
import scipy.linalg

import oracle
from refkit import m, get_l, get_u, get_p, approx_eq, perm_from_ipiv


# ----------------------------------------------------------- reference KATs
def test_lu_square():
    """lu.rs:281-289: l*u == p*a (exact, reference Mul)."""
    a = m([1.0, 2.0, 0.0], [3.0, 6.0, -1.0], [1.0, 2.0, 1.0])
    lu, piv, _ = oracle.lu_new(a)
    assert np.array_equal(oracle.mul(get_l(lu), get_u(lu)), oracle.mul(get_p(piv), a))


def test_lu_m_over_n():
    """lu.rs:291-299."""
    a = m([1.0, 2.0], [3.0, 4.0], [5.0, 6.0])
    lu, piv, _ = oracle.lu_new(a)
    assert np.array_equal(oracle.mul(get_l(lu), get_u(lu)), oracle.mul(get_p(piv), a))


def test_lu_m_under_n():
    """lu.rs:301-309."""
    a = m([1.0, 2.0, 3.0], [4.0, 5.0, 6.0])
    lu, piv, _ = oracle.lu_new(a)
    assert np.array_equal(oracle.mul(get_l(lu), get_u(lu)), oracle.mul(get_p(piv), a))


def test_lu_solve():
    """lu.rs:311-317: solve -> [-1, 3, 3] (approx_eq)."""
    a = m([2.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0])
    lu, piv, _ = oracle.lu_new(a)
    assert approx_eq(oracle.lu_solve(lu, piv, m([1.0], [2.0], [3.0])), m([-1.0], [3.0], [3.0]))


def test_lu_solve_incompatible():
    """lu.rs:319-326: should_panic -> SHAPE."""
    a = m([2.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0])
    lu, piv, _ = oracle.lu_new(a)
    with pytest.raises(oracle.OracleError) as e:
        oracle.lu_solve(lu, piv, m([1.0], [2.0], [3.0], [4.0]))
    assert e.value.status == oracle.SHAPE


def test_lu_solve_singular():
    """lu.rs:328-334: None."""
    lu, piv, _ = oracle.lu_new(m([2.0, 6.0], [1.0, 3.0]))
    assert oracle.lu_solve(lu, piv, m([1.0], [2.0])) is None


def test_lu_is_singular():
    """lu.rs:336-356."""
    assert not oracle.lu_is_non_singular(oracle.lu_new(m([2.0, 6.0], [1.0, 3.0]))[0])
    assert oracle.lu_is_non_singular(oracle.lu_new(m([2.0, 6.0], [1.0, 4.0]))[0])
    assert oracle.lu_is_non_singular(oracle.lu_new(m([4.0, 8.0], [3.0, 4.0]))[0])
    assert not oracle.lu_is_non_singular(oracle.lu_new(m([4.0, 6.0], [2.0, 3.0]))[0])


def test_lu_det():
    """lu.rs:358-367: exact -8.0 and 0.0."""
    lu, _, s = oracle.lu_new(m([4.0, 8.0], [3.0, 4.0]))
    assert oracle.lu_det(lu, s) == -8.0
    lu, _, s = oracle.lu_new(m([4.0, 8.0], [2.0, 4.0]))
    assert oracle.lu_det(lu, s) == 0.0


def test_lu_det_not_square():
    """lu.rs:369-375: should_panic."""
    lu, _, s = oracle.lu_new(m([1.0, 2.0, 3.0], [4.0, 5.0, 6.0]))
    with pytest.raises(oracle.OracleError):
        oracle.lu_det(lu, s)


def test_matrix_det():
    """mod.rs:1506-1510: |det - -96| <= f32::EPSILON; the restatement gives
    exactly -96 with piv [0, 2, 1]."""
    lu, piv, s = oracle.lu_new(m([6.0, -7.0, 10.0], [0.0, 3.0, -1.0], [0.0, 5.0, -7.0]))
    assert oracle.lu_det(lu, s) == -96.0
    assert list(piv) == [0, 2, 1]


def test_matrix_solve():
    """mod.rs:1518-1523: exact [1,1,1] (via .eq), piv [2,1,0]."""
    a = m([1.0, 1.0, 1.0], [1.0, -1.0, 4.0], [2.0, 3.0, -5.0])
    lu, piv, _ = oracle.lu_new(a)
    x = oracle.lu_solve(lu, piv, m([3.0], [4.0], [0.0]))
    assert np.array_equal(x, m([1.0], [1.0], [1.0]))
    assert list(piv) == [2, 1, 0]


def test_matrix_inverse():
    """mod.rs:1527-1533: inverse = solve(id) approx [16,-1,23;0,42,-6;0,30,-18]/96."""
    a = m([6.0, -7.0, 10.0], [0.0, 3.0, -1.0], [0.0, 5.0, -7.0])
    lu, piv, _ = oracle.lu_new(a)
    inv = oracle.lu_solve(lu, piv, np.eye(3))
    assert approx_eq(inv, np.array([16.0, -1.0, 23.0, 0.0, 42.0, -6.0, 0.0, 30.0, -18.0]).reshape(3, 3) / 96.0)


def test_matrix_inverse_singular():
    """mod.rs:1542-1546."""
    lu, piv, _ = oracle.lu_new(m([2.0, 6.0], [1.0, 3.0]))
    assert oracle.lu_solve(lu, piv, np.eye(2)) is None


def test_matrix_is_non_singular():
    """mod.rs:1554-1571."""
    assert not oracle.lu_is_non_singular(oracle.lu_new(m([2.0, 6.0], [1.0, 3.0]))[0])
    assert oracle.lu_is_non_singular(oracle.lu_new(m([2.0, 6.0], [6.0, 3.0]))[0])


def test_is_non_singular_panics_when_wide():
    """lu.rs:174-182 indexes lu[j*n+j] for every j < cols: out of bounds
    (a panic) when rows < cols."""
    lu, _, _ = oracle.lu_new(m([1.0, 2.0, 3.0], [4.0, 5.0, 7.0]))
    with pytest.raises(oracle.OracleError):
        oracle.lu_is_non_singular(lu)


def test_mul():
    """mod.rs:1479-1484 and mmatrix.rs:234-241: [13,16,29,36]."""
    assert oracle.mul(m([1, 2], [3, 4]), m([3, 4], [5, 6])).ravel().tolist() == [13, 16, 29, 36]


def test_mul_incompatible():
    """mod.rs:1486-1492: should_panic."""
    with pytest.raises(oracle.OracleError):
        oracle.mul(m([1, 2], [3, 4]), m([1, 2], [3, 4], [5, 6]))


def test_cholesky_square_pos_def():
    """cholesky.rs:147-153: L data exact, L*L^T == A exact."""
    a = m([4.0, 12.0, -16.0], [12.0, 37.0, -43.0], [-16.0, -43.0, 98.0])
    l, st = oracle.chol_new(a)
    assert st == oracle.OK
    assert l.ravel().tolist() == [2.0, 0.0, 0.0, 6.0, 1.0, 0.0, -8.0, 5.0, 3.0]
    assert np.array_equal(oracle.mul(l, l.T.copy()), a)


def test_cholesky_not_pos_def():
    """cholesky.rs:155-159 (this matrix is also asymmetric: 43 vs -43)."""
    l, st = oracle.chol_new(m([4.0, 12.0, -16.0], [12.0, 37.0, 43.0], [-16.0, 43.0, 98.0]))
    assert l is None


def test_cholesky_not_square():
    """cholesky.rs:161-165."""
    l, st = oracle.chol_new(m([4.0, 12.0, -16.0], [12.0, 37.0, 43.0]))
    assert l is None and st == oracle.NOT_SQUARE


def test_cholesky_solve():
    """cholesky.rs:167-173."""
    l, _ = oracle.chol_new(m([2.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]))
    assert approx_eq(oracle.chol_solve(l, m([1.0], [2.0], [3.0])), m([-1.0], [3.0], [3.0]))


def test_cholesky_solve_incompatible():
    """cholesky.rs:175-182: should_panic."""
    l, _ = oracle.chol_new(m([2.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]))
    with pytest.raises(oracle.OracleError):
        oracle.chol_solve(l, m([1.0], [2.0], [3.0], [4.0]))


# ------------------------------------------------ independent pins
def _lapack_perm(a):
    _, ipiv = scipy.linalg.lu_factor(a, check_finite=False)
    return perm_from_ipiv(ipiv, a.shape[0])


@pytest.mark.parametrize("n,seed", [(8, 1), (64, 2), (256, 3), (257, 4)])
def test_pivots_match_lapack(n, seed):
    """LAPACK dgetrf (idamax: first maximal |.|) uses the same pivot rule;
    on a certified tie-free input the permutations must agree."""
    a = oracle.uniform(seed, 0, (n, n))
    lu, piv, sign, gap = oracle.lu_new(a, want_gap=True)
    assert gap > 1e-9
    assert np.array_equal(piv.astype(np.int64), _lapack_perm(a))


@pytest.mark.parametrize("shape", [(40, 25), (25, 40), (128, 128)])
def test_pa_equals_lu(shape):
    a = oracle.normal(7, 1, shape)
    lu, piv, sign = oracle.lu_new(a)
    l, u = get_l(lu), get_u(lu)
    r = l @ u - a[piv.astype(np.int64)]
    assert np.abs(r).max() < 1e-13 * max(1.0, np.abs(a).max()) * shape[1]
    assert sorted(piv.tolist()) == list(range(shape[0]))
    # sign = parity of the permutation
    p = piv.astype(np.int64)
    seen, parity = np.zeros(len(p), bool), 0
    for i in range(len(p)):
        if not seen[i]:
            j, cyc = i, 0
            while not seen[j]:
                seen[j] = True
                j = p[j]
                cyc += 1
            parity += cyc - 1
    assert sign == (parity % 2 == 0)


def test_lu_2x2_closed_form():
    """2x2 with |c| > |a|: piv [1,0], l = a/c, u22 = b - l*d (single IEEE ops)."""
    a_, b_, c_, d_ = 0.3, 0.7, -1.9, 0.45
    lu, piv, sign = oracle.lu_new(np.array([[a_, b_], [c_, d_]]))
    assert list(piv) == [1, 0] and sign is False
    l21 = a_ / c_
    assert lu[0, 0] == c_ and lu[0, 1] == d_ and lu[1, 0] == l21
    assert lu[1, 1] == b_ - l21 * d_


def test_pivot_tie_goes_to_lowest_row():
    """lu.rs:134 uses a strict '>': equal |.| keeps the earlier row."""
    for a in (m([1.0, 2.0], [1.0, 3.0]), m([1.0, 2.0], [-1.0, 3.0]), m([-2.0, 0.0], [2.0, 1.0])):
        _, piv, sign = oracle.lu_new(a)
        assert list(piv) == [0, 1] and sign is True
    _, piv, _ = oracle.lu_new(m([0.0, 1.0], [0.0, 2.0], [0.0, 3.0]))
    assert list(piv) == [0, 2, 1]   # zero first column: no swap, no division; then |3| wins


def test_nan_pivot_rule():
    """A NaN never wins (|x| > NaN is false) and a NaN in place stays put."""
    _, piv, _ = oracle.lu_new(m([np.nan, 1.0], [2.0, 3.0]))
    assert list(piv) == [0, 1]
    _, piv, _ = oracle.lu_new(m([1.0, 1.0], [np.nan, 3.0], [2.0, 0.0]))
    assert list(piv)[0] == 2


@pytest.mark.parametrize("n", [1, 5, 64, 130])
def test_cholesky_matches_numpy(n):
    a = oracle.spd_matrix(n, 11, 5)
    l, st = oracle.chol_new(a)
    assert st == oracle.OK
    ln = np.linalg.cholesky(a)
    assert np.abs(l - ln).max() < 1e-13 * np.abs(ln).max()
    assert np.all(np.triu(l, 1) == 0.0)


def test_cholesky_failure_kinds():
    a = oracle.spd_matrix(6, 2, 2)
    b = a.copy(); b[4, 1] = np.nextafter(b[4, 1], 10.0)
    assert oracle.chol_new(b)[1] == oracle.NOT_SYMMETRIC
    c = a.copy(); c[3, 3] = -1.0
    assert oracle.chol_new(c)[1] == oracle.NOT_SPD
    d = a.copy(); d[2, 2] = np.nan          # NaN pivot is not rejected (d <= 0 false)
    l, st = oracle.chol_new(d)
    assert st == oracle.OK and np.isnan(l[2, 2])


@pytest.mark.parametrize("shape", [(3, 5, 7), (64, 64, 64), (100, 37, 81)])
def test_mul_matches_numpy(shape):
    mm, kk, pp = shape
    a, b = oracle.uniform(5, 0, (mm, kk)), oracle.uniform(5, 1, (kk, pp))
    c = oracle.mul(a, b)
    assert np.abs(c - a @ b).max() < 4 * kk * 2.2e-16
    ai, bi = np.round(a * 8), np.round(b * 8)        # small integers: exact
    assert np.array_equal(oracle.mul(ai, bi), ai @ bi)


def test_solve_matches_scipy():
    a = oracle.uniform(9, 0, (200, 200))
    b = oracle.uniform(9, 2, (200, 3))
    lu, piv, _ = oracle.lu_new(a)
    x = oracle.lu_solve(lu, piv, b)
    xs = scipy.linalg.solve(a, b)
    assert np.abs(x - xs).max() / np.abs(xs).max() < 1e-11


# ---------------------------------------------------------- generators
def _splitmix_np(seed, stream, count):
    G = np.uint64(0x9E3779B97F4A7C15)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))
    with np.errstate(over="ignore"):
        key = mix(np.uint64(seed) + np.uint64(stream + 1) * G)
        i = np.arange(1, count + 1, dtype=np.uint64)
        return mix(key + i * G)


def test_uniform_generator_bits():
    """The C generator equals an independent numpy SplitMix64 bit for bit."""
    bits = _splitmix_np(123, 4, 1000)
    u = (bits >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    assert np.array_equal(oracle.uniform(123, 4, (1000,)), 2.0 * u - 1.0)
    assert oracle.rng_u64(123, 4, 0) == int(bits[0])


def test_normal_generator():
    bits = _splitmix_np(5, 6, 2000)
    u1 = ((bits[0::2] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (bits[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    ref = np.empty(2000)
    ref[0::2] = r * np.cos(2 * np.pi * u2)
    ref[1::2] = r * np.sin(2 * np.pi * u2)
    got = oracle.normal(5, 6, (2000,))
    assert np.abs(got - ref).max() < 1e-14
    big = oracle.normal(1, 1, (200000,))
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1.0) < 0.01


def test_householder_generator_closed_form():
    """A = U diag(s) V^T: singular values are exactly s, and x* = A^-1 b."""
    n, kappa = 96, 1e4
    a, us, vs, sig = oracle.householder_svd_matrix(n, kappa, 8, seed=3)
    sv = np.linalg.svd(a, compute_uv=False)
    assert np.abs(sv - sig).max() < 1e-13
    assert abs(np.linalg.cond(a) / kappa - 1.0) < 1e-8
    b = oracle.normal(3, 300, (n, 2))
    xs = oracle.householder_exact_solution(us, vs, sig, b)
    assert np.abs(a @ xs - b).max() < 1e-12 * np.abs(xs).max()
    assert np.abs(xs - np.linalg.solve(a, b)).max() / np.abs(xs).max() < kappa * 1e-14


def test_oracle_reproduces_cfg1_fixture():
    """The committed config-1 fixture is the oracle's own output; this keeps
    the oracle from drifting (bit-exact)."""
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cfg1_n256.json")))
    a = oracle.uniform(g["seed"], 0, (256, 256))
    b = oracle.uniform(g["seed"], 2, (256, 1))
    lu, piv, sign = oracle.lu_new(a)
    assert piv.tolist() == g["piv"] and sign == g["pospivsign"]
    assert oracle.lu_solve(lu, piv, b).ravel().tolist() == g["x"]
    assert oracle.lu_det(lu, sign) == g["det"]


# ------------------------------------------------ f32 surface (simd.rs)
def test_oracle_smul_exact_on_integers():
    """simd.rs:189-219 on small-integer matrices: every product and partial
    sum is exact in FP32, so the result equals the exact product; shapes
    with p mod 4 != 0 exercise the scalar tail columns."""
    rng = np.random.default_rng(3)
    for (m, k, p) in [(3, 4, 5), (1, 1, 1), (7, 9, 4), (5, 3, 11)]:
        a = rng.integers(-20, 21, size=(m, k)).astype(np.float32)
        b = rng.integers(-20, 21, size=(k, p)).astype(np.float32)
        assert np.array_equal(oracle.smul(a, b), (a.astype(np.float64) @ b).astype(np.float32))


def test_oracle_smul_rounding_bound_and_shape():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((33, 70)).astype(np.float32)
    b = rng.standard_normal((70, 21)).astype(np.float32)
    c = oracle.smul(a, b).astype(np.float64)
    exact = a.astype(np.float64) @ b.astype(np.float64)
    bound = 70 * 2.0 ** -24 * (np.abs(a).astype(np.float64) @ np.abs(b))
    assert (np.abs(c - exact) <= bound).all()
    # a transposed operand or a dropped term would break the bound
    assert not np.allclose(c, (a.astype(np.float64) @ b[::-1]), atol=1e-3)
    with pytest.raises(oracle.OracleError):
        oracle.smul(a, a)


def test_oracle_sdot():
    x = np.arange(7, dtype=np.float32)
    assert oracle.sdot(x, x) == 91.0                       # 0^2 + ... + 6^2, tail included
    assert oracle.sdot(x[:3], x[:3]) == 5.0                # shorter than one lane group
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal(1001).astype(np.float32), rng.standard_normal(1001).astype(np.float32)
    exact = float(a.astype(np.float64) @ b.astype(np.float64))
    assert abs(oracle.sdot(a, b) - exact) <= 1001 * 2.0 ** -24 * float(np.abs(a) @ np.abs(b))


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (3, 2), (64, 64), (300, 300), (257, 130)])
def test_oracle_lu_colcopy_bit_identical(shape):
    """The contiguous-column variant used to write the n = 8192 fixture does
    the same operations in the same order as the statement of lu.rs:104-168."""
    a = oracle.normal(11, shape[0], shape)
    r1 = oracle.lu_new(a, want_gap=True)
    r2 = oracle.lu_new_colcopy(a, want_gap=True)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    assert r1[2] == r2[2] and r1[3] == r2[3]


# --------------------------------------------------------------- QR (qr.rs)
QR_KATS = [
    m([12.0, -51.0, 4.0], [6.0, 167.0, -68.0], [-4.0, 24.0, -41.0]),   # qr.rs:241-245
    m([1.0, 2.0], [3.0, 4.0], [5.0, 6.0]),                               # qr.rs:248-252
    m([1.0, 2.0, 3.0], [4.0, 5.0, 6.0]),                                 # qr.rs:255-259
]


@pytest.mark.parametrize("k", range(len(QR_KATS)))
def test_qr_kats(k):
    """The reference's three QR tests: (get_q() * get_r()).approx_eq(a)."""
    a = QR_KATS[k]
    qr, rd = oracle.qr_new(a)
    assert approx_eq(oracle.mul(oracle.qr_q(qr, rd), oracle.qr_r(qr, rd)), a)


def test_qr_against_the_closed_form_of_the_first_kat():
    """qr.rs:241's matrix has the textbook factor R = [[-14, -21, 14],
    [0, -175, 70], [0, 0, 35]] with the reference's sign rule (a = -|x| for
    x_00 = 12 > 0, ...): rdiag and the upper part follow exactly."""
    qr, rd = oracle.qr_new(QR_KATS[0])
    assert np.allclose(rd, [-14.0, -175.0, 35.0], rtol=0, atol=1e-12)
    assert np.allclose(np.triu(qr, 1)[[0, 0, 1], [1, 2, 2]], [-21.0, 14.0, 70.0], atol=1e-12)


@pytest.mark.parametrize("shape", [(7, 7), (9, 4), (4, 9), (40, 25)])
def test_qr_invariants_and_lapack(shape):
    """Q R = A, Q's first min(m, n) columns orthonormal, R = LAPACK's R up to
    the sign of each row (the reference fixes the sign by a = -sign(x) |x|),
    the Householder vectors reproduce the reflections."""
    a = oracle.normal(31, shape[0] * 100 + shape[1], shape)
    mm, nn = shape
    qr, rd = oracle.qr_new(a)
    q, r = oracle.qr_q(qr, rd), oracle.qr_r(qr, rd)
    k = min(mm, nn)
    assert np.abs(q @ r - a).max() < 1e-13 * max(1.0, np.abs(a).max()) * max(shape)
    assert np.abs(q[:, :k].T @ q[:, :k] - np.eye(k)).max() < 1e-13 * max(shape)
    assert np.all(q[:, k:] == 0.0)
    r_lapack = scipy.linalg.qr(a, mode="economic")[1]
    assert np.allclose(np.abs(r[:k]), np.abs(r_lapack[:k]), atol=1e-12 * np.abs(a).max() * max(shape))
    # H_c applied by hand: I + u u^T / (a u_c)
    h = np.eye(mm)
    for c in range(k):
        u = np.zeros(mm)
        u[c:] = qr[c:, c]
        h = h @ (np.eye(mm) + np.outer(u, u) / (rd[c] * qr[c, c]))
    assert np.abs(h[:, :k] - q[:, :k]).max() < 1e-12 * max(shape)


def test_qr_zero_column_and_rank():
    """A zero column gives a = 0 and is left as it was (qr.rs:62); the factor
    is then not of full rank (qr.rs:108-115), and is_full_rank panics (index
    out of bounds) when n > m."""
    a = oracle.normal(5, 1, (6, 4))
    a[:, 2] = 0.0
    qr, rd = oracle.qr_new(a)
    # column 2 is zero only above the diagonal after the first two
    # reflections, so rdiag[2] is the norm of a vector that stays zero
    assert rd[2] == 0.0 and not oracle.qr_is_full_rank(qr, rd)
    with pytest.raises(oracle.OracleError):
        oracle.qr_is_full_rank(*oracle.qr_new(oracle.normal(5, 2, (3, 5))))


def test_qr_solve_least_squares_and_the_literal_defect():
    """The corrected solve is the least-squares solution (numpy lstsq); the
    reference's literal formula (qr.rs:211-222, JAMA's normalisation applied
    to unnormalised vectors) is not: its residual on qr.rs:241's matrix is
    in the thousands."""
    a = QR_KATS[0]
    b = m([1.0], [2.0], [3.0])
    qr, rd = oracle.qr_new(a)
    x = oracle.qr_solve(qr, rd, b)
    assert np.abs(a @ x - b).max() < 1e-13
    xl = oracle.qr_solve_literal(qr, rd, b)
    assert np.abs(a @ xl - b).max() > 1e3
    tall = oracle.normal(8, 3, (30, 7))
    bt = oracle.normal(8, 4, (30, 2))
    xt = oracle.qr_solve(*oracle.qr_new(tall), bt)
    assert np.allclose(xt, np.linalg.lstsq(tall, bt, rcond=None)[0], atol=1e-12)
    with pytest.raises(oracle.OracleError):
        oracle.qr_solve_literal(*oracle.qr_new(tall), bt)        # Matrix::new assert, m > n


def test_pinverse_kat_and_numpy():
    """mod.rs:1549-1552: a.pinverse() * a ~ id(2); and numpy's pinv for a
    full-column-rank matrix."""
    a = m([1.0, 2.0], [3.0, 4.0], [5.0, 6.0])
    assert approx_eq(oracle.mul(oracle.pinverse(a), a), np.eye(2))
    b = oracle.normal(9, 9, (20, 6))
    assert np.allclose(oracle.pinverse(b), np.linalg.pinv(b), atol=1e-12)
