"""CPU-side checks of the C ABI: the library loads and exports exactly what
include/tara_hip.h declares.  No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tara_amd", "libtara_hip.so")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tara_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libtara_hip.so not built (python -m tara_amd.build)")
    return ctypes.CDLL(LIB)


def test_header_symbols_exported(lib):
    names = declared("tara_hip.h")
    assert len(names) >= 10
    exported = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", exported), f"{n} declared but not exported"
        getattr(lib, n)


def test_version_and_error_string(lib):
    assert lib.tara_version() >= 100
    lib.tara_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.tara_last_error(), bytes)


def test_argument_errors_without_gpu(lib):
    """Argument validation happens before any device call."""
    lib.tara_last_error.restype = ctypes.c_char_p
    rc = lib.tara_derive2csh(64, 64, 30, ctypes.c_double(0.01), None, None)
    assert rc == -1
    rc = lib.tara_csh_create(64, 64, ctypes.c_double(0.01), ctypes.c_double(1e-4), 7, ctypes.byref(ctypes.c_void_p()))
    assert rc == -1
    assert b"scheme" in lib.tara_last_error()


def test_host_philox_matches_known_answers():
    """tara_philox4x32_10 is host code: it runs without a GPU."""
    import tara_amd as T
    from test_oracle import PHILOX_KAT
    for ctr, key, want in PHILOX_KAT:
        assert T.philox4x32_10(ctr, key) == want
