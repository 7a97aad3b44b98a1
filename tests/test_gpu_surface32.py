"""GPU tests of the f32 surface (src/matrix/simd.rs:6-219) through the C ABI.

simd_mul runs on the FP32 MFMA GEMM, whose products and sums round
differently from the reference's sequential f32x4 loop, so it is held to the
FP32 rounding bound against the oracle (k 2^-24 |A||B| elementwise) and is
exact on small-integer inputs; the elementwise operations are single IEEE
operations and must match bit for bit; simd_dot must be within the FP32
rounding bound of the oracle's dot product (which corrects the reference's
tail defect)."""
import ctypes
import zlib

import numpy as np
import pytest

import oracle
import la_hip

pytestmark = pytest.mark.gpu
_fp = ctypes.POINTER(ctypes.c_float)
EW = {"scale": 0, "add": 1, "sub": 2, "mul": 3, "div": 4, "neg": 5}


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def sgemm(ctx, a, b):
    a, b = _f(a), _f(b)
    c = np.empty((a.shape[0], b.shape[1]), dtype=np.float32)
    st = ctx.lib.la_sgemm_rm(ctx.h, a.shape[0], a.shape[1], b.shape[0], b.shape[1],
                             a.ctypes.data_as(_fp), b.ctypes.data_as(_fp), c.ctypes.data_as(_fp))
    return st, c


def elementwise(ctx, op, a, b=None, s=0.0):
    a = _f(a)
    bb = _f(b) if b is not None else a
    out = np.empty_like(a)
    st = ctx.lib.la_selementwise_rm(ctx.h, EW[op], a.shape[0], a.shape[1], bb.shape[0],
                                    bb.shape[1], a.ctypes.data_as(_fp),
                                    bb.ctypes.data_as(_fp) if b is not None else None,
                                    ctypes.c_float(s), out.ctypes.data_as(_fp))
    return st, out


@pytest.fixture(scope="module", autouse=True)
def _sig(ctx):
    L = ctx.lib
    i64 = ctypes.c_int64
    L.la_sgemm_rm.argtypes = [ctypes.c_void_p, i64, i64, i64, i64, _fp, _fp, _fp]
    L.la_selementwise_rm.argtypes = [ctypes.c_void_p, ctypes.c_int, i64, i64, i64, i64, _fp, _fp,
                                     ctypes.c_float, _fp]
    L.la_sdot_rm.argtypes = [ctypes.c_void_p, i64, i64, i64, i64, _fp, _fp, _fp]


@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 1, 1), (129, 65, 257), (256, 256, 256),
                                   (300, 17, 3)])
def test_simd_mul_vs_oracle(ctx, shape):
    m, k, p = shape
    rng = np.random.default_rng(m * 7 + k)
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, p)).astype(np.float32)
    st, c = sgemm(ctx, a, b)
    assert st == la_hip.LA_OK
    ref = oracle.smul(a, b).astype(np.float64)
    bound = 2 * k * 2.0 ** -24 * (np.abs(a).astype(np.float64) @ np.abs(b)) + 1e-30
    assert (np.abs(c - ref) <= bound).all()


def test_simd_mul_exact_on_integers(ctx):
    rng = np.random.default_rng(9)
    a = rng.integers(-30, 31, size=(70, 40)).astype(np.float32)
    b = rng.integers(-30, 31, size=(40, 33)).astype(np.float32)
    st, c = sgemm(ctx, a, b)
    assert st == la_hip.LA_OK and np.array_equal(c, oracle.smul(a, b))


def test_simd_mul_shape_error(ctx):
    st, _ = sgemm(ctx, np.ones((2, 3)), np.ones((2, 3)))
    assert st == la_hip.LA_ERR_SHAPE


@pytest.mark.parametrize("shape", [(1, 1), (1, 3), (5, 7), (64, 65)])
@pytest.mark.parametrize("op", ["scale", "add", "sub", "mul", "div", "neg"])
def test_elementwise_bit_exact(ctx, shape, op):
    rng = np.random.default_rng(zlib.crc32(repr((shape, op)).encode()))
    a = rng.standard_normal(shape).astype(np.float32)
    b = rng.standard_normal(shape).astype(np.float32)
    s = np.float32(1.7)
    want = {"scale": s * a, "add": a + b, "sub": a - b, "mul": a * b, "div": a / b,
            "neg": -a}[op]
    st, out = elementwise(ctx, op, a, None if op in ("scale", "neg") else b, float(s))
    assert st == la_hip.LA_OK
    assert np.array_equal(out.view(np.uint32), want.astype(np.float32).view(np.uint32))


def test_elementwise_shape_error(ctx):
    st, _ = elementwise(ctx, "add", np.ones((2, 3)), np.ones((3, 2)))
    assert st == la_hip.LA_ERR_SHAPE


@pytest.mark.parametrize("n", [1, 3, 4, 7, 1000, 100003])
def test_simd_dot(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    out = np.empty(1, dtype=np.float32)
    st = ctx.lib.la_sdot_rm(ctx.h, n, 1, n, 1, a.ctypes.data_as(_fp), b.ctypes.data_as(_fp),
                            out.ctypes.data_as(_fp))
    assert st == la_hip.LA_OK
    ref = oracle.sdot(a, b)
    bound = 2 * n * 2.0 ** -24 * float(np.abs(a).astype(np.float64) @ np.abs(b)) + 1e-30
    assert abs(float(out[0]) - ref) <= bound
    st = ctx.lib.la_sdot_rm(ctx.h, n, 2, n, 2, a.ctypes.data_as(_fp), b.ctypes.data_as(_fp),
                            out.ctypes.data_as(_fp))
    assert st == la_hip.LA_ERR_SHAPE
