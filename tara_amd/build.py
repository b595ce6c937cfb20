"""Builds libtara_hip.so (gfx950) in-tree from tara_amd/csrc/*.hip, and the
Fortran drop-in wrappers (libtara_fortran.so) with amdflang.

    python -m tara_amd.build            # compile everything
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtara_hip.so")
FLIB = os.path.join(HERE, "libtara_fortran.so")
ARCH = os.environ.get("TARA_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build_hip(jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(OBJ, s[:-4] + ".o") for s in srcs]

    def one(pair):
        src, obj = pair
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])

    with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    probe = os.path.join(HERE, "calib", "copy_probe")
    _run([HIPCC] + HIP_FLAGS + [probe + ".hip", "-o", probe + ".tmp"])
    os.replace(probe + ".tmp", probe)
    return LIB


def build_shim():
    """libtara_dfftw.so: FFTW's Fortran entry points over libtara_hip.so (host C++)."""
    out = os.path.join(HERE, "libtara_dfftw.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
          os.path.join(HERE, "shim", "dfftw_shim.cpp"), "-o", tmp, "-L" + HERE, "-ltara_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_gnucompat():
    """libtara_gnucompat.so: GNU Fortran's RAND / SRAND for the kept mains (host C++)."""
    out = os.path.join(HERE, "libtara_gnucompat.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(HERE, "shim", "gnu_compat.cpp"),
          "-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran():
    """ISO_C_BINDING wrappers keeping the reference's subroutine names
    (libtara_fortran.so), and the Fortran test drivers under tests/fortran/."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    fdir = os.path.join(HERE, "fortran")
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(fdir) if f.endswith(".f90"))
    order = sorted(srcs, key=lambda s: (0 if s.endswith("_mod.f90") else 1, s))  # modules first
    objs = []
    for s in order:
        o = os.path.join(OBJ, s[:-4] + ".fo")
        _run([fc, "-O2", "-fPIC", "-J" + OBJ, "-c", os.path.join(fdir, s), "-o", o])
        objs.append(o)
    tmp = FLIB + ".tmp"
    _run([fc, "-shared", "-o", tmp] + objs + ["-L" + HERE, "-ltara_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    tdir = os.path.join(ROOT, "tests", "fortran")
    for s in sorted(f for f in os.listdir(tdir) if f.endswith(".f90")):
        exe = os.path.join(tdir, s[:-4])
        libs = ["-ltara_dfftw"] if s.startswith("drive_dfftw") else ["-ltara_fortran"]
        _run([fc, "-O2", "-J" + OBJ, os.path.join(tdir, s), "-o", exe + ".tmp", "-L" + HERE] + libs +
             ["-ltara_hip", "-Wl,-rpath,$ORIGIN/../../tara_amd"])
        os.replace(exe + ".tmp", exe)
    return FLIB


def main():
    build_hip()
    build_shim()
    build_gnucompat()
    build_fortran()
    print(LIB)


if __name__ == "__main__":
    main()
