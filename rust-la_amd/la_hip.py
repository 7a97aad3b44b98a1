"""ctypes binding of liblahip.so (C ABI declared in include/la_hip.h).

Host-pointer calls take and return numpy arrays and mirror the reference's
drop-in surface (Mul, mmul, LUDecomposition::new/solve,
CholeskyDecomposition::new/solve).  Device-pointer calls take torch CUDA
tensors (torch is only the device-memory and stream plumbing here).

There is no fallback: if the library is missing this module raises on
import of the library, never computes on the CPU.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblahip.so")

LA_OK = 0
LA_ERR_SHAPE = 1
LA_SINGULAR = 2
LA_NOT_SQUARE = 3
LA_NOT_SYMMETRIC = 4
LA_NOT_SPD = 5
LA_ERR_HIP = 6
LA_ERR_NOMEM = 7
LA_NOT_CONVERGED = 8
LA_ERR_ARG = 9
LA_ERR_TIMEOUT = 10
LA_ERR_IO = 11

PREC_FP64, PREC_FP32, PREC_BF16, PREC_FP16 = 0, 1, 2, 3
IR_RBT = 0x100      # option bit for gesv_ir_dev: butterfly transform, no pivoting
IR_GMRES = 0x200    # option bit for gesv_ir_dev: GMRES-preconditioned corrections
IR_TOURNAMENT = 0x400  # option bit for gesv_ir_dev: tournament-pivoted (CALU) FP32 factor
IR_METHOD_PIVOTED, IR_METHOD_RBT, IR_METHOD_RBT_THEN_PIVOTED = 0, 1, 2
IR_METHOD_TOURNAMENT, IR_METHOD_TOURNAMENT_THEN_PIVOTED = 3, 4
IR_METHOD_NAMES = ["pivoted", "butterfly", "butterfly then pivoted", "tournament",
                   "tournament then pivoted"]

KERNEL_CLASSES = ["gemm64", "gemm32", "gemm16", "panel", "trsv", "resid", "trsm", "rowmove",
                  "batch", "pivsel"]

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


class LaError(RuntimeError):
    """A non-OK status from the library (LA_ERR_SHAPE is the reference's panic)."""

    def __init__(self, status, msg=""):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class IrReport(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("prec", ctypes.c_int32), ("method", ctypes.c_int32),
                ("backward_error", ctypes.c_double),
                ("initial_backward_error", ctypes.c_double),
                ("anorm_inf", ctypes.c_double), ("factor_ms", ctypes.c_double),
                ("refine_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("perturbed_pivots", ctypes.c_int32), ("inner_iterations", ctypes.c_int32)]


_lib = None

# Live contexts.  Teardown is explicit and ordered: close() releases a
# context's plans, then the context; the atexit hook does the same for
# whatever is still open while the HIP runtime is fully alive, and no
# destructor calls into the library once the interpreter is finalising.
_live_contexts = weakref.WeakSet()


def _close_all():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


atexit.register(_close_all)


def load():
    """Load liblahip.so; raises if it was not built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    L.la_status_string.restype = ctypes.c_char_p
    L.la_ctx_last_error.restype = ctypes.c_char_p
    L.la_ctx_last_error.argtypes = [_vp]
    L.la_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.la_ctx_destroy.argtypes = [_vp]
    L.la_ctx_synchronize.argtypes = [_vp]
    L.la_ctx_stream.argtypes = [_vp]
    L.la_ctx_stream.restype = _vp
    L.la_ctx_set_stream.argtypes = [_vp, _vp]
    L.la_device_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _i64p,
                                 ctypes.POINTER(ctypes.c_int)]
    L.la_ctx_set_workspace_limit.argtypes = [_vp, _i64]
    L.la_dgemm_rm.argtypes = [_vp, _i64, _i64, _i64, _i64, _dp, _dp, _dp]
    L.la_dmmul_rm.argtypes = [_vp] + [_i64] * 6 + [_dp, _dp, _dp]
    L.la_dgetrf_rm.argtypes = [_vp, _i64, _i64, _dp, _dp, _u64p, ctypes.POINTER(ctypes.c_int)]
    L.la_dgetrs_rm.argtypes = [_vp, _i64, _i64, _dp, _u64p, _i64, _i64, _dp, _dp]
    L.la_dgesv_rm.argtypes = [_vp, _i64, _i64, _dp, _i64, _i64, _dp, _dp]
    L.la_dpotrf_rm.argtypes = [_vp, _i64, _i64, _dp, _dp]
    L.la_dpotrs_rm.argtypes = [_vp, _i64, _dp, _i64, _i64, _dp, _dp]
    L.la_dgemm_dev.argtypes = [_vp, _i64, _i64, _i64, ctypes.c_double, _vp, _i64, _vp, _i64,
                               ctypes.c_double, _vp, _i64]
    L.la_dgetrf_dev.argtypes = [_vp, _i64, _i64, _vp, _i64, _vp, _vp]
    L.la_dgetrf_noalloc_dev.argtypes = [_vp, _i64, _i64, _vp, _i64, _vp, _vp]
    L.la_dgetrs_dev.argtypes = [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]
    L.la_dpotrf_dev.argtypes = [_vp, _i64, _vp, _i64, _vp]
    L.la_dpotrs_dev.argtypes = [_vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64]
    L.la_ctx_set_profiling.argtypes = [_vp, ctypes.c_int]
    L.la_ctx_reset_kernel_stats.argtypes = [_vp]
    L.la_ctx_kernel_stats.argtypes = [_vp, ctypes.c_int, _i64p, _dp, _dp, _dp]
    L.la_fill_uniform_dev.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint64, _i64, _vp]
    L.la_fill_normal_dev.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint64, _i64, _vp]
    for name, args in _SIGNATURES.items():
        getattr(L, name).argtypes = args      # a missing export fails here, loudly
    _lib = L
    return L


_SIGNATURES = {
    "la_dgesv_ir_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, ctypes.c_int,
                        ctypes.c_double, ctypes.c_int, ctypes.POINTER(IrReport)],
    "la_dgesv_ir_rm": [_vp, _i64, _i64, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double,
                       ctypes.c_int, ctypes.POINTER(IrReport)],
    "la_gen_svd_matrix_dev": [_vp, _i64, ctypes.c_double, ctypes.c_int, ctypes.c_uint64,
                              ctypes.c_uint64, ctypes.c_uint64, _vp, _i64, _i64,
                              ctypes.c_uint64, _vp, _vp],
    "la_backward_error_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                              ctypes.POINTER(ctypes.c_double)],
    "la_batch_plan_create": [_vp, ctypes.c_int, _i64, _i32p, _i32p, ctypes.c_int,
                             ctypes.POINTER(_vp)],
    "la_batch_plan_destroy": [_vp],
    "la_batch_plan_describe": [_vp],
    "la_dpotrf_potrs_batched_dev": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "la_dgesv_ir_batched_dev": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                ctypes.c_int, _vp, _vp, _vp],
    "la_gen_spd_batch_dev": [_vp, _i64, _vp, _vp, ctypes.c_uint64, ctypes.c_uint64, _vp, _vp],
    "la_gen_svd_batch_dev": [_vp, _i64, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                             ctypes.c_uint64, ctypes.c_uint64, _vp, _vp, _vp],
    "la_dinverse_rm": [_vp, _i64, _i64, _dp, _dp],
    "la_dgeqrf_rm": [_vp, _i64, _i64, _dp, _dp, _dp],
    "la_dqr_q_rm": [_vp, _i64, _i64, _dp, _dp, _dp],
    "la_dqrs_rm": [_vp, _i64, _i64, _dp, _dp, _i64, _i64, _dp, _dp],
    "la_dpinverse_rm": [_vp, _i64, _i64, _dp, _dp],
    "la_dgeqrf_dev": [_vp, _i64, _i64, _vp, _i64, _vp],
    "la_dqr_q_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64],
    "la_dqrs_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64],
    "la_dinverse_dev": [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp],
    "la_sgemm_dev": [_vp, _i64, _i64, _i64, ctypes.c_float, _vp, _i64, _vp, _i64,
                     ctypes.c_float, _vp, _i64],
    "la_autotune": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p],
    "la_autotune_load": [_vp, ctypes.c_char_p],
}


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class Context:
    """One la_ctx (device, stream, workspaces)."""

    def __init__(self, device=0):
        # PyTorch-ROCm bundles its own HIP runtime; when both live in one
        # process, torch's must be initialised first or torch then reports
        # no device.  Initialise it here when torch is importable.
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
        self.lib = load()
        h = _vp()
        self.h = None
        self._plans = weakref.WeakSet()
        st = self.lib.la_ctx_create(device, ctypes.byref(h))
        if st:
            raise LaError(st, "la_ctx_create")
        self.h = h
        _live_contexts.add(self)

    def close(self):
        """Synchronise, destroy this context's batch plans, then the context."""
        for p in list(self._plans):
            p.close()
        if self.h:
            self.lib.la_ctx_destroy(self.h)
            self.h = None
        _live_contexts.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        # at interpreter shutdown the atexit hook has already closed every
        # context; never call into HIP from a finalising interpreter
        if sys is None or sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != LA_OK:
            msg = self.lib.la_ctx_last_error(self.h).decode(errors="replace")
            raise LaError(st, f"{what}: {msg}")

    def sync(self):
        self._check(self.lib.la_ctx_synchronize(self.h), "synchronize")

    def stream(self):
        return self.lib.la_ctx_stream(self.h)

    def set_stream(self, hip_stream):
        """Run the library on an existing stream (e.g. torch's current
        stream, so torch-side copies and library kernels stay ordered)."""
        self._check(self.lib.la_ctx_set_stream(self.h, ctypes.c_void_p(hip_stream)), "set_stream")

    def use_torch_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream().cuda_stream)

    def set_profiling(self, on):
        self._check(self.lib.la_ctx_set_profiling(self.h, 1 if on else 0), "set_profiling")

    def reset_kernel_stats(self):
        self._check(self.lib.la_ctx_reset_kernel_stats(self.h), "reset_kernel_stats")

    def kernel_stats(self):
        """{class: (calls, total_ms, algorithmic flop, algorithmic bytes)}"""
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            c, ms, fl, by = _i64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            self._check(self.lib.la_ctx_kernel_stats(self.h, i, ctypes.byref(c), ctypes.byref(ms),
                                                     ctypes.byref(fl), ctypes.byref(by)),
                        "kernel_stats")
            if c.value:
                out[name] = (c.value, ms.value, fl.value, by.value)
        return out

    def device_info(self):
        cus, hbm, clk = ctypes.c_int(), _i64(), ctypes.c_int()
        self.lib.la_device_info(self.h, ctypes.byref(cus), ctypes.byref(hbm), ctypes.byref(clk))
        return cus.value, hbm.value, clk.value

    # ---- host (drop-in) entry points ------------------------------------
    def mul(self, a, b):
        a, b = _c(a), _c(b)
        c = np.empty((a.shape[0], b.shape[1]))
        self._check(self.lib.la_dgemm_rm(self.h, a.shape[0], a.shape[1], b.shape[0], b.shape[1],
                                         _d(a), _d(b), _d(c)), "dgemm_rm")
        return c

    def mmul(self, a, b, dst):
        a, b = _c(a), _c(b)
        assert dst.flags.c_contiguous and dst.dtype == np.float64
        self._check(self.lib.la_dmmul_rm(self.h, a.shape[0], a.shape[1], b.shape[0], b.shape[1],
                                         dst.shape[0], dst.shape[1], _d(a), _d(b), _d(dst)),
                    "dmmul_rm")
        return dst

    def lu_new(self, a):
        a = _c(a)
        m, n = a.shape
        lu = np.empty_like(a)
        piv = np.empty(m, dtype=np.uint64)
        sign = ctypes.c_int()
        self._check(self.lib.la_dgetrf_rm(self.h, m, n, _d(a), _d(lu), piv.ctypes.data_as(_u64p),
                                          ctypes.byref(sign)), "dgetrf_rm")
        return lu, piv, bool(sign.value)

    def lu_solve(self, lu, piv, b):
        """Returns None where the reference returns None (singular)."""
        lu, b = _c(lu), _c(b)
        if b.ndim == 1:
            b = b.reshape(-1, 1)
        piv = np.ascontiguousarray(piv, dtype=np.uint64)
        x = np.empty((lu.shape[0], b.shape[1]))
        st = self.lib.la_dgetrs_rm(self.h, lu.shape[0], lu.shape[1], _d(lu),
                                   piv.ctypes.data_as(_u64p), b.shape[0], b.shape[1], _d(b), _d(x))
        if st == LA_SINGULAR:
            return None
        self._check(st, "dgetrs_rm")
        return x

    def solve(self, a, b):
        a, b = _c(a), _c(b)
        if b.ndim == 1:
            b = b.reshape(-1, 1)
        x = np.empty((a.shape[0], b.shape[1]))
        st = self.lib.la_dgesv_rm(self.h, a.shape[0], a.shape[1], _d(a), b.shape[0], b.shape[1],
                                  _d(b), _d(x))
        if st == LA_SINGULAR:
            return None
        self._check(st, "dgesv_rm")
        return x

    def inverse(self, a):
        """Matrix::inverse (mod.rs:1034-1037); None where the reference returns None."""
        a = _c(a)
        out = np.empty_like(a)
        st = self.lib.la_dinverse_rm(self.h, a.shape[0], a.shape[1], _d(a), _d(out))
        if st == LA_SINGULAR:
            return None
        self._check(st, "dinverse_rm")
        return out

    def qr_new(self, a):
        """QRDecomposition::new (qr.rs:26-106) -> (qr, rdiag)."""
        a = _c(a)
        m, n = a.shape
        qr = np.empty_like(a)
        rdiag = np.empty(min(m, n))
        self._check(self.lib.la_dgeqrf_rm(self.h, m, n, _d(a), _d(qr), _d(rdiag)), "dgeqrf_rm")
        return qr, rdiag

    def qr_q(self, qr, rdiag):
        """QRDecomposition::get_q (qr.rs:150-194), m x m."""
        qr, rdiag = _c(qr), _c(rdiag)
        m, n = qr.shape
        q = np.empty((m, m))
        self._check(self.lib.la_dqr_q_rm(self.h, m, n, _d(qr), _d(rdiag), _d(q)), "dqr_q_rm")
        return q

    def qr_solve(self, qr, rdiag, b):
        """QRDecomposition::solve (qr.rs:196-237) as a least-squares solve:
        (x or None, status)."""
        qr, rdiag, b = _c(qr), _c(rdiag), _c(b)
        if b.ndim == 1:
            b = b.reshape(-1, 1)
        m, n = qr.shape
        x = np.empty((n, b.shape[1]))
        st = self.lib.la_dqrs_rm(self.h, m, n, _d(qr), _d(rdiag), b.shape[0], b.shape[1], _d(b),
                                 _d(x))
        return (x if st == LA_OK else None), st

    def pinverse(self, a):
        """Matrix::pinverse (mod.rs:1049-1057): (n x m or None, status)."""
        a = _c(a)
        m, n = a.shape
        p = np.empty((n, m))
        st = self.lib.la_dpinverse_rm(self.h, m, n, _d(a), _d(p))
        return (p if st == LA_OK else None), st

    def geqrf_dev(self, A, rdiag):
        m, n = A.shape
        self._check(self.lib.la_dgeqrf_dev(self.h, m, n, _ptr(A), A.stride(0), _ptr(rdiag)),
                    "dgeqrf_dev")

    def qr_q_dev(self, QR, rdiag, Q):
        m, n = QR.shape
        self._check(self.lib.la_dqr_q_dev(self.h, m, n, _ptr(QR), QR.stride(0), _ptr(rdiag),
                                          _ptr(Q), Q.stride(0)), "dqr_q_dev")

    def qrs_dev(self, QR, rdiag, B, X):
        """Returns the status (LA_OK, LA_SINGULAR or LA_ERR_SHAPE)."""
        m, n = QR.shape
        return self.lib.la_dqrs_dev(self.h, m, n, _ptr(QR), QR.stride(0), _ptr(rdiag),
                                    B.shape[1], _ptr(B), B.stride(0), _ptr(X), X.stride(0))

    def chol_new(self, a):
        """Returns (L or None, status)."""
        a = _c(a)
        rows, cols = a.shape
        l = np.empty((rows, rows)) if rows == cols else np.empty((1, 1))
        st = self.lib.la_dpotrf_rm(self.h, rows, cols, _d(a), _d(l))
        if st in (LA_NOT_SQUARE, LA_NOT_SYMMETRIC, LA_NOT_SPD):
            return None, st
        self._check(st, "dpotrf_rm")
        return l, st

    def chol_solve(self, l, b):
        l, b = _c(l), _c(b)
        if b.ndim == 1:
            b = b.reshape(-1, 1)
        x = np.empty_like(b)
        self._check(self.lib.la_dpotrs_rm(self.h, l.shape[0], _d(l), b.shape[0], b.shape[1],
                                          _d(b), _d(x)), "dpotrs_rm")
        return x

    # ---- device entry points (torch tensors) -----------------------------
    def gemm_dev(self, A, B, C, alpha=1.0, beta=0.0):
        m, k = A.shape
        n = B.shape[1]
        self._check(self.lib.la_dgemm_dev(self.h, m, n, k, alpha, _ptr(A), A.stride(0), _ptr(B),
                                          B.stride(0), beta, _ptr(C), C.stride(0)), "dgemm_dev")

    def getrf_dev(self, A, piv, info):
        m, n = A.shape
        self._check(self.lib.la_dgetrf_dev(self.h, m, n, _ptr(A), A.stride(0), _ptr(piv),
                                           _ptr(info)), "dgetrf_dev")

    def getrf_noalloc_dev(self, A, piv, info):
        """la_dgetrf_dev's factorisation with the library's fixed scratch
        (no allocation; min(m, n) <= 65536; overlapping calls race)."""
        m, n = A.shape
        self._check(self.lib.la_dgetrf_noalloc_dev(self.h, m, n, _ptr(A), A.stride(0),
                                                   _ptr(piv), _ptr(info)), "dgetrf_noalloc_dev")

    def getrs_dev(self, LU, piv, B, X, status=None):
        m, n = LU.shape
        self._check(self.lib.la_dgetrs_dev(self.h, m, n, _ptr(LU), LU.stride(0), _ptr(piv),
                                           B.shape[1], _ptr(B), B.stride(0), _ptr(X),
                                           X.stride(0), _ptr(status)), "dgetrs_dev")

    def inverse_dev(self, A, piv, X, info):
        """la_dinverse_dev: A becomes its packed LU, X its inverse."""
        n = A.shape[0]
        self._check(self.lib.la_dinverse_dev(self.h, n, _ptr(A), A.stride(0), _ptr(piv), _ptr(X),
                                             X.stride(0), _ptr(info)), "dinverse_dev")

    def sgemm_dev(self, A, B, C, alpha=1.0, beta=0.0):
        """la_sgemm_dev (Matrix<f32>::simd_mul's product on FP32 MFMA)."""
        m, k = A.shape
        n = B.shape[1]
        self._check(self.lib.la_sgemm_dev(self.h, m, n, k, alpha, _ptr(A), A.stride(0), _ptr(B),
                                          B.stride(0), beta, _ptr(C), C.stride(0)), "sgemm_dev")

    def potrf_dev(self, A, status):
        n = A.shape[0]
        self._check(self.lib.la_dpotrf_dev(self.h, n, _ptr(A), A.stride(0), _ptr(status)),
                    "dpotrf_dev")

    def potrs_dev(self, L, B, X):
        n = L.shape[0]
        self._check(self.lib.la_dpotrs_dev(self.h, n, _ptr(L), L.stride(0), B.shape[1], _ptr(B),
                                           B.stride(0), _ptr(X), X.stride(0)), "dpotrs_dev")

    def fill_uniform_dev(self, seed, stream, out):
        self._check(self.lib.la_fill_uniform_dev(self.h, seed, stream, out.numel(), _ptr(out)),
                    "fill_uniform")

    def fill_normal_dev(self, seed, stream, out):
        self._check(self.lib.la_fill_normal_dev(self.h, seed, stream, out.numel(), _ptr(out)),
                    "fill_normal")

    def gesv_ir_dev(self, A, B, X, prec, tol, max_iter):
        rep = IrReport()
        n = A.shape[0]
        st = self.lib.la_dgesv_ir_dev(self.h, n, B.shape[1], _ptr(A), A.stride(0), _ptr(B),
                                      B.stride(0), _ptr(X), X.stride(0), prec, tol, max_iter,
                                      ctypes.byref(rep))
        # non-convergence and a singular factor are outcomes carried by the report
        if st not in (LA_OK, LA_NOT_CONVERGED, LA_SINGULAR):
            self._check(st, "dgesv_ir_dev")
        return rep

    def gen_svd_matrix_dev(self, n, kappa, nrefl, seed, A, u_stream=100, v_stream=200,
                           b=None, xstar=None, b_stream=300):
        nrhs = b.shape[1] if b is not None else 0
        self._check(self.lib.la_gen_svd_matrix_dev(self.h, n, kappa, nrefl, seed, u_stream,
                                                   v_stream, _ptr(A), A.stride(0), nrhs,
                                                   b_stream, _ptr(b), _ptr(xstar)),
                    "gen_svd_matrix")

    # ---- batched uneven solvers ------------------------------------------
    def batch_plan(self, kind, sizes, nrhs=None, prec=0):
        return BatchPlan(self, kind, sizes, nrhs, prec)

    def potrf_potrs_batched_dev(self, plan, offA, offB, A, L, B, X, status):
        self._check(self.lib.la_dpotrf_potrs_batched_dev(self.h, plan.h, _ptr(offA), _ptr(offB),
                                                         _ptr(A), _ptr(L), _ptr(B), _ptr(X),
                                                         _ptr(status)), "dpotrf_potrs_batched")

    def gen_spd_batch_dev(self, d_n, d_offA, seed, stream_base, A, G):
        self._check(self.lib.la_gen_spd_batch_dev(self.h, d_n.numel(), _ptr(d_n), _ptr(d_offA),
                                                  seed, stream_base, _ptr(A), _ptr(G)),
                    "gen_spd_batch")

    def gesv_ir_batched_dev(self, plan, offA, offB, A, B, X, tol, max_iter, status, iters,
                            berr):
        """Batched LU + FP64 refinement (plan kind 1); per-problem status,
        refinement steps and backward error land in the device arrays."""
        self._check(self.lib.la_dgesv_ir_batched_dev(self.h, plan.h, _ptr(offA), _ptr(offB),
                                                     _ptr(A), _ptr(B), _ptr(X), float(tol),
                                                     int(max_iter), _ptr(status), _ptr(iters),
                                                     _ptr(berr)), "dgesv_ir_batched")

    def gen_svd_batch_dev(self, d_n, d_nrhs, offA, offB, kappa, nrefl, seed, stream_base, A, B,
                          Xstar):
        self._check(self.lib.la_gen_svd_batch_dev(self.h, d_n.numel(), _ptr(d_n), _ptr(d_nrhs),
                                                  _ptr(offA), _ptr(offB), float(kappa),
                                                  int(nrefl), seed, stream_base, _ptr(A), _ptr(B),
                                                  _ptr(Xstar)), "gen_svd_batch")

    def autotune(self, kind, nmin, nmax, cache_path=None):
        self._check(self.lib.la_autotune(self.h, kind, nmin, nmax,
                                         cache_path.encode() if cache_path else None), "autotune")

    def autotune_load(self, cache_path):
        self._check(self.lib.la_autotune_load(self.h, cache_path.encode()), "autotune_load")

    def backward_error_dev(self, A, X, B):
        eta = ctypes.c_double()
        self._check(self.lib.la_backward_error_dev(self.h, A.shape[0], B.shape[1], _ptr(A),
                                                   A.stride(0), _ptr(X), X.stride(0), _ptr(B),
                                                   B.stride(0), ctypes.byref(eta)),
                    "backward_error")
        return eta.value


class BatchPlan:
    """la_batch_plan: sorted problem order, device queue, captured graph."""

    def __init__(self, ctx, kind, sizes, nrhs=None, prec=0):
        self.ctx = ctx
        self.h = None
        n = np.ascontiguousarray(sizes, dtype=np.int32)
        r = np.ascontiguousarray(nrhs, dtype=np.int32) if nrhs is not None else None
        h = _vp()
        st = ctx.lib.la_batch_plan_create(ctx.h, kind, len(n), n.ctypes.data_as(_i32p),
                                          r.ctypes.data_as(_i32p) if r is not None else None,
                                          prec, ctypes.byref(h))
        ctx._check(st, "batch_plan_create")
        self.h = h
        ctx._plans.add(self)
        ctx.lib.la_batch_plan_describe.restype = ctypes.c_char_p

    def describe(self):
        return self.ctx.lib.la_batch_plan_describe(self.h).decode()

    def close(self):
        if self.h:
            self.ctx.lib.la_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys is None or sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
