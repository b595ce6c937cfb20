"""The Fortran boundary: the reference's routine names (derive2csh, ab2csh)
resolved to tara_amd/fortran/csh_dropin.f90 -> C ABI -> HIP, driven by a
Fortran time loop (tests/fortran/drive_csh.f90, the loop of
src/main2csh.f95:114-126), against the reference trajectory."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_csh")


def _write(path, w):
    nh, ny = w.shape
    with open(path, "wb") as f:
        np.array([2 * (nh - 1), ny], dtype="<i4").tofile(f)
        np.asfortranarray(w).T.astype("<c16").tofile(f)  # column-major order on disk


def _read(path):
    with open(path, "rb") as f:
        nx, ny = np.fromfile(f, dtype="<i4", count=2)
        w = np.fromfile(f, dtype="<c16", count=(nx // 2 + 1) * ny)
    return np.asfortranarray(w.reshape(ny, nx // 2 + 1).T)


def test_fortran_driver_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "tara_amd", "libtara_fortran.so")],
                         capture_output=True, text=True).stdout
    for sym in ("derive2csh_", "ab2csh_"):
        assert f" T {sym}" in out


def test_io_roundtrip(tmp_path):
    w = np.asfortranarray(np.arange(33 * 64).reshape(33, 64) * (1 + 2j))
    _write(tmp_path / "a.bin", w)
    assert np.array_equal(_read(tmp_path / "a.bin"), w)


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["asrun", "intent"])
def test_fortran_dropin_config1(golden, tmp_path, tag):
    assert os.path.exists(DRIVER), "tests/fortran/drive_csh not built (run __graft_entry__.build())"
    _write(tmp_path / "in.bin", golden[f"c1_{tag}_w0"])
    print(f"fortran drop-in {tag}", flush=True)
    subprocess.run([DRIVER, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "2000", "0.01", "1.0d-4"],
                   check=True, timeout=120)
    assert rel_l2(_read(tmp_path / "out.bin"), golden[f"c1_{tag}_w2000"]) < 1e-10
