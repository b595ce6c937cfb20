"""The drop-in boundary at link level: the reference's own main programs
(src/main.f95 and every driver it dispatches to, with lib/iniparser), compiled
unchanged, link against tara_amd's libraries alone, with every hot-path
routine, every FFTW entry point and GNU RAND/SRAND resolved there and no FFTW
library or reference routine in the link (tests/fortran/Makefile.mains).
Needs the reference checkout and amdflang; no GPU."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
TARA = os.path.join(ROOT, "tara_amd")
pytestmark = pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "src")) and shutil.which("amdflang")),
                                reason="reference checkout or amdflang not present")

HOT_PATH = ["derive2csh_", "ab2csh_", "derive2css_", "ab2css_", "derive2coh_", "ab2coh_", "derive2com_", "ab2com_",
            "derive3com_", "ab3com_"]


def _symbols(path, undefined):
    out = subprocess.run(["nm", "-D", "--undefined-only" if undefined else "--defined-only", path],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_reference_mains_link_against_tara_only(tmp_path):
    for lib in ("libtara_fortran.so", "libtara_dfftw.so", "libtara_gnucompat.so", "libtara_hip.so"):
        assert os.path.exists(os.path.join(TARA, lib)), f"{lib} not built"
    out = tmp_path / "mains"
    r = subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "tests", "fortran", "Makefile.mains"),
                        f"OUT={out}", f"REF={REF}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    exe = str(out / "tara")
    need = _symbols(exe, undefined=True)
    # every hot-path routine, FFTW entry point and RAND/SRAND the mains call is
    # resolved by the tara_amd libraries
    provided = set()
    for lib in ("libtara_fortran.so", "libtara_dfftw.so", "libtara_gnucompat.so"):
        provided |= _symbols(os.path.join(TARA, lib), undefined=False)
    for sym in HOT_PATH + ["rand_", "srand_"]:
        assert sym in need, f"{sym} not called by the mains"
        assert sym in provided, f"{sym} not provided by tara_amd"
    dfftw = {s for s in need if s.startswith("dfftw_")}
    assert len(dfftw) >= 10 and dfftw <= provided, sorted(dfftw - provided)
    # no FFTW library and no reference routine in the link
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "fftw" not in ldd.replace("tara_dfftw", "")
    objs = os.listdir(out)
    assert not any(o.startswith(("derive", "ab2", "ab3", "rk4", "euler", "pc")) for o in objs), objs
    # the config-1 program with flang-safe snapshot OPENs links the same way
    snap = str(out / "tara_snap")
    assert _symbols(snap, undefined=True) >= set(HOT_PATH[:2] + ["rand_"])
    assert "fftw" not in subprocess.run(["ldd", snap], capture_output=True, text=True).stdout.replace("tara_dfftw", "")
