"""ctypes front end of the CPU oracle (la_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() check and by
bench.py's cpu_baseline leg, never by the library.  Each wrapper names the
reference routine it restates (file:line in xasmx/rust-la).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblaoracle.so")
_lib = None

OK, SHAPE, SINGULAR, NOT_SQUARE, NOT_SYMMETRIC, NOT_SPD = 0, 1, 2, 3, 4, 5

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64 = ctypes.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_mul.argtypes = [_i64, _i64, _i64, _i64, _dp, _dp, _dp]
        L.orc_mmul.argtypes = [_i64] * 6 + [_dp, _dp, _dp]
        L.orc_lu_new.argtypes = [_i64, _i64, _dp, _dp, _u64p, ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_lu_new_colcopy.argtypes = L.orc_lu_new.argtypes
        L.orc_lu_is_non_singular.argtypes = [_i64, _i64, _dp]
        L.orc_lu_det.argtypes = [_i64, _i64, _dp, ctypes.c_int, _dp]
        L.orc_lu_solve.argtypes = [_i64, _i64, _dp, _u64p, _i64, _i64, _dp, _dp]
        L.orc_chol_new.argtypes = [_i64, _i64, _dp, _dp, ctypes.POINTER(_i64)]
        L.orc_chol_solve.argtypes = [_i64, _dp, _i64, _i64, _dp, _dp]
        L.orc_rng_u64.argtypes = [ctypes.c_uint64] * 3
        L.orc_rng_u64.restype = ctypes.c_uint64
        L.orc_fill_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _i64, _dp]
        L.orc_fill_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _i64, _dp]
        _fp = ctypes.POINTER(ctypes.c_float)
        L.orc_smul.argtypes = [_i64, _i64, _i64, _i64, _fp, _fp, _fp]
        L.orc_sdot.argtypes = [_i64, _fp, _fp]
        L.orc_sdot.restype = ctypes.c_float
        L.orc_qr_new.argtypes = [_i64, _i64, _dp, _dp, _dp]
        L.orc_qr_is_full_rank.argtypes = [_i64, _i64, _dp]
        L.orc_qr_q.argtypes = [_i64, _i64, _dp, _dp, _dp]
        L.orc_qr_solve.argtypes = [_i64, _i64, _dp, _dp, _i64, _i64, _dp, _dp]
        L.orc_qr_solve_literal.argtypes = L.orc_qr_solve.argtypes
        L.orc_pinverse.argtypes = [_i64, _i64, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleError(Exception):
    def __init__(self, status):
        super().__init__(status)
        self.status = status


def mul(a, b):
    """impl Mul for &Matrix<f64> (src/matrix/mod.rs:957-980)."""
    a, b = _c(a), _c(b)
    c = np.empty((a.shape[0], b.shape[1]))
    st = lib().orc_mul(a.shape[0], a.shape[1], b.shape[0], b.shape[1], _d(a), _d(b), _d(c))
    if st:
        raise OracleError(st)
    return c


def lu_new(a, want_gap=False):
    """LUDecomposition::new (src/decomp/lu.rs:104-168) -> (lu, piv, pospivsign[, gap])."""
    a = _c(a)
    m, n = a.shape
    lu = np.empty_like(a)
    piv = np.empty(m, dtype=np.uint64)
    sign = ctypes.c_int()
    gap = ctypes.c_double()
    st = lib().orc_lu_new(m, n, _d(a), _d(lu), piv.ctypes.data_as(_u64p), ctypes.byref(sign),
                          ctypes.byref(gap))
    if st:
        raise OracleError(st)
    out = (lu, piv, bool(sign.value))
    return out + (gap.value,) if want_gap else out


def lu_new_colcopy(a, want_gap=False):
    """lu_new with the column walk made contiguous (bit-identical results;
    for writing large fixtures)."""
    a = _c(a)
    m, n = a.shape
    lu = np.empty_like(a)
    piv = np.empty(m, dtype=np.uint64)
    sign = ctypes.c_int()
    gap = ctypes.c_double()
    st = lib().orc_lu_new_colcopy(m, n, _d(a), _d(lu), piv.ctypes.data_as(_u64p),
                                  ctypes.byref(sign), ctypes.byref(gap))
    if st:
        raise OracleError(st)
    out = (lu, piv, bool(sign.value))
    return out + (gap.value,) if want_gap else out


def lu_is_non_singular(lu):
    """lu.rs:174-182; raises OracleError(SHAPE) where the reference panics."""
    lu = _c(lu)
    r = lib().orc_lu_is_non_singular(lu.shape[0], lu.shape[1], _d(lu))
    if r < 0:
        raise OracleError(SHAPE)
    return bool(r)


def lu_det(lu, sign):
    """lu.rs:224-232."""
    lu = _c(lu)
    d = ctypes.c_double()
    st = lib().orc_lu_det(lu.shape[0], lu.shape[1], _d(lu), int(sign), ctypes.byref(d))
    if st:
        raise OracleError(st)
    return d.value


def lu_solve(lu, piv, b):
    """LUDecomposition::solve (lu.rs:237-278); None for a singular factor."""
    lu, b = _c(lu), _c(b)
    if b.ndim == 1:
        b = b.reshape(-1, 1)
    m, n = lu.shape
    x = np.empty((m, b.shape[1]))
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    st = lib().orc_lu_solve(m, n, _d(lu), piv.ctypes.data_as(_u64p), b.shape[0], b.shape[1],
                            _d(b), _d(x))
    if st == SINGULAR:
        return None
    if st:
        raise OracleError(st)
    return x


def chol_new(a):
    """CholeskyDecomposition::new (src/decomp/cholesky.rs:56-110); None on failure.
    Returns (L, status)."""
    a = _c(a)
    rows, cols = a.shape
    l = np.zeros((rows, rows)) if rows == cols else None
    fr = _i64(-1)
    st = lib().orc_chol_new(rows, cols, _d(a), _d(l) if l is not None else None, ctypes.byref(fr))
    return (l if st == OK else None), st


def chol_solve(l, b):
    """CholeskyDecomposition::solve (cholesky.rs:116-144)."""
    l, b = _c(l), _c(b)
    if b.ndim == 1:
        b = b.reshape(-1, 1)
    x = np.empty_like(b)
    st = lib().orc_chol_solve(l.shape[0], _d(l), b.shape[0], b.shape[1], _d(b), _d(x))
    if st:
        raise OracleError(st)
    return x


def smul(a, b):
    """Matrix<f32>::simd_mul (src/matrix/simd.rs:189-219), FP32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    c = np.empty((a.shape[0], b.shape[1]), dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    st = lib().orc_smul(a.shape[0], a.shape[1], b.shape[0], b.shape[1], a.ctypes.data_as(fp),
                        b.ctypes.data_as(fp), c.ctypes.data_as(fp))
    if st:
        raise OracleError(st)
    return c


def sdot(a, b):
    """simd_dot (simd.rs:166-187) with the tail adding a[i]*b[i]."""
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    fp = ctypes.POINTER(ctypes.c_float)
    return float(lib().orc_sdot(a.size, a.ctypes.data_as(fp), b.ctypes.data_as(fp)))


def uniform(seed, stream, shape):
    out = np.empty(int(np.prod(shape)))
    lib().orc_fill_uniform(seed, stream, out.size, _d(out))
    return out.reshape(shape)


def normal(seed, stream, shape):
    out = np.empty(int(np.prod(shape)))
    lib().orc_fill_normal(seed, stream, out.size, _d(out))
    return out.reshape(shape)


def rng_u64(seed, stream, i):
    return lib().orc_rng_u64(seed, stream, i)


# ---- test-matrix generators built on the seeded streams -------------------

def householder_svd_matrix(n, kappa, nrefl, seed, u_stream=100, v_stream=200):
    """A = U diag(sigma) V^T with U = H_1...H_nrefl, V = H'_1...H'_nrefl,
    H_i = I - 2 v v^T / (v^T v), v_i = normal(seed, u_stream + i) (resp.
    v_stream + i); sigma_k = kappa^(-k/(n-1)).  Returns (A, U-vectors,
    V-vectors, sigma)."""
    sig = kappa ** (-np.arange(n) / max(n - 1, 1))
    us = [normal(seed, u_stream + i, (n,)) for i in range(nrefl)]
    vs = [normal(seed, v_stream + i, (n,)) for i in range(nrefl)]
    m = np.diag(sig)
    for v in reversed(us):           # M <- H_i M, i = nrefl..1
        m = m - np.outer(2.0 * v / (v @ v), v @ m)
    for v in reversed(vs):           # M <- M H'_i, i = nrefl..1
        m = m - np.outer(m @ v, 2.0 * v / (v @ v))
    return m, us, vs, sig


def householder_exact_solution(us, vs, sig, b):
    """x* = V diag(1/sigma) U^T b through the reflectors."""
    y = np.array(b, dtype=np.float64, copy=True)
    if y.ndim == 1:
        y = y.reshape(-1, 1)
    # U^T b = H_nrefl ... H_2 H_1 b : apply H_1 first
    for v in us:
        y = y - np.outer(2.0 * v / (v @ v), v @ y)
    y = y / sig[:, None]
    # V y = H'_1 ... H'_nrefl y : apply H'_nrefl first
    for v in reversed(vs):
        y = y - np.outer(2.0 * v / (v @ v), v @ y)
    return y


def spd_matrix(n, seed, stream):
    """A = G G^T / n + I, G standard normal (stream)."""
    g = normal(seed, stream, (n, n))
    return g @ g.T / n + np.eye(n)


def qr_new(a):
    """QRDecomposition::new (src/decomp/qr.rs:26-106) -> (qr, rdiag)."""
    a = _c(a)
    m, n = a.shape
    qr = np.empty_like(a)
    rdiag = np.empty(min(m, n))
    st = lib().orc_qr_new(m, n, _d(a), _d(qr), _d(rdiag))
    if st:
        raise OracleError(st)
    return qr, rdiag


def qr_is_full_rank(qr, rdiag):
    """QRDecomposition::is_full_rank (qr.rs:108-115); raises on the reference's
    out-of-bounds panic (n > m)."""
    m, n = qr.shape
    r = lib().orc_qr_is_full_rank(m, n, _d(_c(rdiag)))
    if r < 0:
        raise OracleError(SHAPE)
    return bool(r)


def qr_h(qr):
    """QRDecomposition::get_h (qr.rs:117-132): the lower trapezoid of qr."""
    return np.tril(qr)


def qr_r(qr, rdiag):
    """QRDecomposition::get_r (qr.rs:134-148): m x n, rdiag on the diagonal."""
    r = np.triu(qr, 1)
    k = len(rdiag)
    r[np.arange(k), np.arange(k)] = rdiag
    return r


def qr_q(qr, rdiag):
    """QRDecomposition::get_q (qr.rs:150-194): m x m."""
    qr = _c(qr)
    m, n = qr.shape
    q = np.empty((m, m))
    lib().orc_qr_q(m, n, _d(qr), _d(_c(rdiag)), _d(q))
    return q


def _qr_solve(fn, qr, rdiag, b):
    qr, b = _c(qr), _c(b)
    if b.ndim == 1:
        b = b.reshape(-1, 1)
    m, n = qr.shape
    x = np.empty((max(m, n), b.shape[1]))
    st = fn(m, n, _d(qr), _d(_c(rdiag)), b.shape[0], b.shape[1], _d(b), _d(x))
    return st, x


def qr_solve(qr, rdiag, b):
    """The least-squares solve QRDecomposition::solve means (qr.rs:196-237 with
    the reflections of qr.rs:76-104): n x nx, or the status on failure."""
    st, x = _qr_solve(lib().orc_qr_solve, qr, rdiag, b)
    if st:
        raise OracleError(st)
    return x[:qr.shape[1]]


def qr_solve_literal(qr, rdiag, b):
    """QRDecomposition::solve exactly as written (qr.rs:196-237; m == n)."""
    st, x = _qr_solve(lib().orc_qr_solve_literal, qr, rdiag, b)
    if st:
        raise OracleError(st)
    return x[:qr.shape[1]]


def pinverse(a):
    """Matrix::pinverse (src/matrix/mod.rs:1049-1057): n x m."""
    a = _c(a)
    m, n = a.shape
    p = np.empty((n, m))
    st = lib().orc_pinverse(m, n, _d(a), _d(p))
    if st:
        raise OracleError(st)
    return p
