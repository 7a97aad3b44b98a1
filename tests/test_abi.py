"""CPU-only checks of the C ABI library: it was built, it loads, and it
exports every entry point include/la_hip.h declares.  No compute calls."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rust-la_amd", "liblahip.so")
HDR = os.path.join(ROOT, "include", "la_hip.h")


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(la_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("la_dgemm_rm", "la_dgetrf_rm", "la_dgetrs_rm", "la_dpotrf_rm", "la_dpotrs_rm"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "liblahip.so not built"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_status_strings_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.la_status_string.restype = ctypes.c_char_p
    assert b"panic" in lib.la_status_string(1)
    assert b"None" in lib.la_status_string(2)


def test_oracle_is_not_linked_into_the_library():
    """The product must not route through the CPU oracle."""
    lib = ctypes.CDLL(LIB)
    assert not hasattr(lib, "orc_lu_new")
    data = open(LIB, "rb").read()
    assert b"orc_lu_new" not in data and b"la_oracle" not in data
