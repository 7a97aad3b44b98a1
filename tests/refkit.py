"""Host-side helpers shared by the tests: the reference's small matrix
accessors restated in numpy (lu.rs:184-220, mod.rs:416-438, 757-759) and the
reference's approx_eq (src/approxeq.rs:34-47: absolute 1e-6)."""
import numpy as np


def m(*rows):
    """m!(a, b; c, d) literal (src/macros.rs:39-42)."""
    return np.array(rows, dtype=np.float64)


def get_l(lu):
    """lu.rs:184-202: m x min(m,n), unit diagonal."""
    mm, nn = lu.shape
    k = min(mm, nn)
    l = np.zeros((mm, k))
    for i in range(mm):
        for j in range(k):
            l[i, j] = lu[i, j] if i > j else (1.0 if i == j else 0.0)
    return l


def get_u(lu):
    """lu.rs:204-215: min(m,n) x n."""
    mm, nn = lu.shape
    k = min(mm, nn)
    u = np.zeros((k, nn))
    for i in range(k):
        for j in range(nn):
            u[i, j] = lu[i, j] if i <= j else 0.0
    return u


def get_p(piv):
    """lu.rs:217-220: id(len,len).permute_rows(piv)."""
    n = len(piv)
    return np.eye(n)[np.asarray(piv, dtype=np.int64)]


def approx_eq(a, b):
    """ApproxEq for f64 (approxeq.rs:45): |a-b| < 1e-6 elementwise."""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all(np.abs(a - b) < 1e-6))


def perm_from_ipiv(ipiv, m):
    """LAPACK swap list -> permutation vector (row i of PA = row p[i] of A)."""
    p = np.arange(m)
    for k, t in enumerate(ipiv):
        p[k], p[t] = p[t], p[k]
    return p
