"""CPU-side checks of the C ABI: the library loads and exports exactly what
include/tara_hip.h declares; host-only entry points (I/O, the counter-based
generator) run here.  No GPU compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
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


def test_checkpoint_roundtrip(tmp_path):
    """Host I/O: runs without a GPU."""
    from tara_amd import snapio
    rng = np.random.default_rng(0)
    f = [np.asfortranarray(rng.standard_normal((9, 16)) + 1j * rng.standard_normal((9, 16))) for _ in range(3)]
    snapio.write_checkpoint(tmp_path / "c.bin", f, 16, step=42, time=0.5)
    g, h = snapio.read_checkpoint(tmp_path / "c.bin")
    assert h == {"nx": 16, "ny": 16, "nz": 1, "step": 42, "time": 0.5}
    assert all(np.array_equal(a, b) for a, b in zip(f, g))
    f3 = [np.asfortranarray(rng.standard_normal((5, 8, 4)) + 0j) for _ in range(2)]
    snapio.write_checkpoint(tmp_path / "d.bin", f3, 8)
    g3, h3 = snapio.read_checkpoint(tmp_path / "d.bin")
    assert h3["nz"] == 4 and all(np.array_equal(a, b) for a, b in zip(f3, g3))


def test_pair_store_fp32_and_bf16(tmp_path):
    from tara_amd import snapio
    rng = np.random.default_rng(1)
    a = [np.asfortranarray(rng.standard_normal((16, 8))) for _ in range(6)]
    for dt, tol in ((snapio.FP32, 1e-7), (snapio.BF16, 2.0 ** -8)):
        w = snapio.PairWriter(tmp_path / f"p{dt}.bin", 16, 8, dt, dt_pair=0.05, nu=1e-3)
        for k in range(3):
            w.append(a[2 * k], a[2 * k + 1])
        w.close()
        x, y, h = snapio.read_pairs(tmp_path / f"p{dt}.bin")
        assert h["npairs"] == 3 and h["dt_pair"] == 0.05 and x.shape == (3, 16, 8)
        for k in range(3):
            assert np.max(np.abs(x[k] - a[2 * k]) / (np.abs(a[2 * k]) + 1e-30)) <= tol
            assert np.max(np.abs(y[k] - a[2 * k + 1]) / (np.abs(a[2 * k + 1]) + 1e-30)) <= tol
    # bf16 rounding is to nearest even: 1 + 2^-8 (a tie) rounds down to 1, 1 + 3 * 2^-8 up
    w = snapio.PairWriter(tmp_path / "t.bin", 2, 1, snapio.BF16)
    w.append(np.array([[1 + 2.0 ** -8], [1 + 3 * 2.0 ** -8]]), np.zeros((2, 1)))
    w.close()
    x, _, _ = snapio.read_pairs(tmp_path / "t.bin")
    assert x[0, 0, 0] == 1.0 and x[0, 1, 0] == 1 + 2.0 ** -6


def test_fort_text_snapshot(tmp_path):
    from tara_amd import snapio
    om = np.asfortranarray(np.arange(8.0).reshape(4, 2))
    snapio.write_fort(tmp_path / "fort.00000100", om, 2 * om, 3 * om, 2 * np.pi, 2 * np.pi)
    rows = np.loadtxt(tmp_path / "fort.00000100")
    assert rows.shape == (8, 5)
    # i outer, j inner (src/main2csh.f95:170-178)
    assert np.allclose(rows[1], [0.0, np.pi, om[0, 1], 2 * om[0, 1], 3 * om[0, 1]])
    assert np.allclose(rows[2, :2], [np.pi / 2, 0.0])


def test_fortran_file_ic_hook(tmp_path):
    """tara_read_state2d / tara_write_state2d (tara_amd/fortran/tara_hip_mod.f90),
    the file initial condition of a kept main, round-trip a checkpoint written
    from Python (host I/O only)."""
    import subprocess

    from tara_amd import snapio

    exe = os.path.join(ROOT, "tests", "fortran", "drive_ckpt")
    if not os.path.exists(exe):
        pytest.skip("tests/fortran/drive_ckpt not built (amdflang absent)")
    nx, ny = 32, 16
    rng = np.random.default_rng(4)
    w = np.asfortranarray(rng.standard_normal((nx // 2 + 1, ny)) + 1j * rng.standard_normal((nx // 2 + 1, ny)))
    snapio.write_checkpoint(tmp_path / "in.ckpt", [w], nx, step=7, time=1.25)
    subprocess.run([exe, str(tmp_path / "in.ckpt"), str(tmp_path / "out.ckpt"), str(nx), str(ny)], check=True,
                   timeout=60)
    fields, h = snapio.read_checkpoint(tmp_path / "out.ckpt")
    assert h["step"] == 8 and h["time"] == 1.75 and (h["nx"], h["ny"], h["nz"]) == (nx, ny, 1)
    assert np.array_equal(fields[0], 2.0 * w)
