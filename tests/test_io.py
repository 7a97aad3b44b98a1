"""CPU tests of the two data formats on either side of the dense path:
read_csv (src/util.rs:9-35) with the f64 item parser, and the Debug
rendering of a Matrix<f64> (src/matrix/mod.rs:815-840).  Both are host code
in the library; these calls touch no GPU.

The reference holds no tests or fixtures for either, so the expectations
are pinned by the reference's own control flow (which inputs panic, through
unwrap() or assert!, and how a row is laid out) and by the documented
behaviour of Rust's `str::parse::<f64>` and `{:?}` for f64 (shortest
round-trip digits; exponent form outside [1e-4, 1e16)).  The checks use
independent means: Python's float repr (also shortest round-trip, same
thresholds) for digits, hand-written strings for layout."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rust-la_amd", "liblahip.so")
LA_OK, LA_ERR_SHAPE, LA_ERR_ARG, LA_ERR_IO = 0, 1, 9, 11


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(LIB)
    i64p = ctypes.POINTER(ctypes.c_int64)
    dp = ctypes.POINTER(ctypes.c_double)
    L.la_csv_shape.argtypes = [ctypes.c_char_p, i64p, i64p]
    L.la_read_csv_rm.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, dp]
    L.la_format_debug.argtypes = [ctypes.c_int64, ctypes.c_int64, dp, ctypes.c_char_p,
                                  ctypes.c_int64]
    L.la_format_debug.restype = ctypes.c_int64
    return L


def read_csv(lib, path):
    r, c = ctypes.c_int64(), ctypes.c_int64()
    st = lib.la_csv_shape(str(path).encode(), ctypes.byref(r), ctypes.byref(c))
    if st:
        return st, None
    out = np.empty((r.value, c.value))
    st = lib.la_read_csv_rm(str(path).encode(), r.value, c.value,
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return st, out


def write(tmp_path, text, name="m.csv"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return p


def test_read_csv_basic(lib, tmp_path):
    st, m = read_csv(lib, write(tmp_path, "1,2,3\n4.5, -6e2 ,  7\n"))
    assert st == LA_OK
    assert m.tolist() == [[1.0, 2.0, 3.0], [4.5, -600.0, 7.0]]


def test_read_csv_crlf_and_no_final_newline(lib, tmp_path):
    st, m = read_csv(lib, write(tmp_path, "0.1,0.2\r\n0.3,1e-310"))
    assert st == LA_OK
    assert m.ravel().tolist() == [0.1, 0.2, 0.3, 1e-310]


def test_read_csv_correct_rounding(lib, tmp_path):
    # halfway and long literals: correctly rounded, like Rust's parser
    items = ["9007199254740993", "0.1000000000000000055511151231257827",
             "2.2250738585072011e-308", "1.7976931348623157e308", "-0.0", ".5", "5.", "+3"]
    st, m = read_csv(lib, write(tmp_path, ",".join(items) + "\n"))
    assert st == LA_OK
    assert [float(x) for x in m.ravel()] == [float(x) for x in items]
    assert np.signbit(m[0, 4])


def test_read_csv_special_values(lib, tmp_path):
    st, m = read_csv(lib, write(tmp_path, "inf,-inf,NaN,Infinity,nan\n"))
    assert st == LA_OK
    v = m.ravel()
    assert v[0] == np.inf and v[1] == -np.inf and np.isnan(v[2]) and v[3] == np.inf
    assert np.isnan(v[4])


@pytest.mark.parametrize("text", ["1,2\n3,x\n", "1,,2\n", "1,2\n\n3,4\n", "0x10\n", "1_000\n",
                                  "1e\n", ".\n", "1 2\n", "--1\n", "1e5.5\n"])
def test_read_csv_unparseable_item_is_io_error(lib, tmp_path, text):
    """util.rs:19 parser(item).unwrap(): an item that is not an f64 literal panics."""
    st, _ = read_csv(lib, write(tmp_path, text))
    assert st == LA_ERR_IO


def test_read_csv_ragged_and_empty_are_shape_errors(lib, tmp_path):
    st, _ = read_csv(lib, write(tmp_path, "1,2\n3\n"))          # util.rs:25-27
    assert st == LA_ERR_SHAPE
    st, _ = read_csv(lib, write(tmp_path, ""))                  # util.rs:33
    assert st == LA_ERR_SHAPE


def test_read_csv_missing_file(lib, tmp_path):
    st, _ = read_csv(lib, tmp_path / "nope.csv")
    assert st == LA_ERR_IO


def test_read_csv_shape_mismatch_is_arg_error(lib, tmp_path):
    p = write(tmp_path, "1,2\n3,4\n")
    out = np.empty((3, 2))
    assert lib.la_read_csv_rm(str(p).encode(), 3, 2,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == LA_ERR_ARG


def debug(lib, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    dp = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    need = lib.la_format_debug(a.shape[0], a.shape[1], dp, None, 0)
    buf = ctypes.create_string_buffer(int(need) + 1)
    assert lib.la_format_debug(a.shape[0], a.shape[1], dp, buf, need + 1) == need
    return buf.value.decode()


def rust_f64_debug(v):
    """Rust {:?} for f64 from Python's repr (both shortest round-trip with the
    same plain/exponent thresholds); only the exponent syntax differs."""
    if np.isnan(v):
        return "NaN"
    if np.isinf(v):
        return "inf" if v > 0 else "-inf"
    r = repr(float(v))
    if "e" in r:
        mant, exp = r.split("e")
        if mant.endswith(".0"):
            mant = mant[:-2]
        return f"{mant}e{int(exp)}"
    return r


@pytest.mark.parametrize("v,s", [(1.0, "1.0"), (0.1, "0.1"), (-0.0, "-0.0"), (0.0, "0.0"),
                                 (1e-4, "0.0001"), (1e-5, "1e-5"), (1.5e-7, "1.5e-7"),
                                 (1e15, "1000000000000000.0"), (1e16, "1e16"),
                                 (123456789012345680.0, "1.2345678901234568e17"),
                                 (1.7976931348623157e308, "1.7976931348623157e308"),
                                 (5e-324, "5e-324"), (2.5, "2.5"), (-33.125, "-33.125"),
                                 (float("nan"), "NaN"), (float("inf"), "inf"),
                                 (float("-inf"), "-inf")])
def test_debug_element_format(lib, v, s):
    assert debug(lib, np.array([[v]])) == f"\n| {s} |\n"
    assert rust_f64_debug(v) == s


def test_debug_matrix_layout(lib):
    """mod.rs:815-840: widths from the widest element, one extra blank."""
    a = np.array([[1.0, 2.5], [-30.25, 4.0]])
    assert debug(lib, a) == "\n|    1.0    2.5 |\n| -30.25    4.0 |\n"


def test_debug_matches_reference_rules_on_random_values(lib):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((7, 5)) * 10.0 ** rng.integers(-8, 20, size=(7, 5))
    items = [[rust_f64_debug(v) for v in row] for row in a]
    w = max(len(s) for row in items for s in row)
    expect = "\n" + "".join("|" + "".join(" " * (1 + w - len(s)) + s for s in row) + " |\n"
                            for row in items)
    assert debug(lib, a) == expect
