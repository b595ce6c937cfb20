"""The reference's own 3D program run end to end through tara_amd: the
dispatcher src/main.f95 -> tara3dMHDOMP (src/main3com.f95), compiled unchanged
and linked against tara_amd only (tests/fortran/_mains/tara: derive3com /
ab3com through the Fortran wrappers to the HIP kernels, and the main's own
21 diagnostic FFTW calls per step through the dfftw shim), against the same
objects linked against the reference's routines and FFTW 3.3.9
(tests/fortran/_mains/tara_ref, built from the reference sources by
oracle/Makefile).  Both write output/Energy.dat once per step (pressure,
kinetic and total energy, B^2, vorticity moments, helicities and div B,
src/main3com.f95:1024-1071); the runs agree to the FP64 parity bar.

Both binaries are built by __graft_entry__.build() where the reference
checkout is present.  The reference allocates its AB2 history uninitialised
(src/main3com.f95 allocates d_*_dt_old without setting it); both runs set
glibc's mmap threshold to one page so that allocations come from fresh,
zeroed pages, the zero history both assume.

(The 2D mains link too, tests/test_link_mains.py, but flang's runtime stops
them at their first text snapshot: they re-open an already connected unit with
STATUS='unknown', e.g. src/main2csh.f95:177, which gfortran tolerates.)"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

MAINS = os.path.join(ROOT, "tests", "fortran", "_mains")
TARA, TARA_REF = os.path.join(MAINS, "tara"), os.path.join(MAINS, "tara_ref")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(TARA) and os.path.exists(TARA_REF)),
                                 reason="reference mains not linked (no reference checkout at build)")]

# the 3D configuration's physics on a 16^3 grid: dim = 3, openmp, mhd
# (tara3dMHDOMP), box 2 pi (input 2.0), dt = 1e-3, 20 loop trips
# (time_max = 19.5 dt), nu = 0.01 (the main's "resistivity" key feeds mu)
INI = """[dimension]
dim = 3
[architecture]
arch = cpu
[parallel]
mode = openmp
[type]
domain = mhd
[problem]
hydrotype = bare
[particle]
Np = 10
[grid]
Nx = {n}
Ny = {n}
Nz = {n}
[length]
Lx = 2.0d0
Ly = 2.0d0
Lz = 2.0d0
[resistivity]
nu = 0.010d0
eta = 0.0d0
[time]
time_min = 0.0d0
time_max = {tmax}d0
dt = 0.0010d0
[initialProfile]
W0 = 25.0d0
m0 = 2.0d0
d = 3.0d0/128.0d0
[initialB]
B0 = 0.0d0
"""
COLUMNS = ["time", "Pressure", "Energy", "T_Energy", "B_Field", "C1", "C2", "Hf", "Hm", "HT", "div_B_Tot"]


def _run(exe, where, n, steps):
    where.mkdir()
    (where / "output").mkdir()
    (where / "input.ini").write_text(INI.format(n=n, tmax=repr((steps - 0.5) * 1e-3)))
    env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="4096")
    r = subprocess.run([exe, "input.ini"], cwd=where, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (exe, r.stdout[-2000:], r.stderr[-2000:])
    text = (where / "output" / "Energy.dat").read_text().replace("D", "E")
    vals = np.array([float(v) for v in text.split()])
    assert vals.size == steps * len(COLUMNS), (exe, vals.size)
    return vals.reshape(steps, len(COLUMNS))


@pytest.mark.parametrize("n,steps", [(16, 20)])
def test_reference_main3com_through_tara_vs_reference_program(tmp_path, n, steps):
    print(f"main3com {n}^3 x {steps}: reference program (CPU)", flush=True)
    want = _run(TARA_REF, tmp_path / "ref", n, steps)
    print(f"main3com {n}^3 x {steps}: linked against tara_amd (GPU)", flush=True)
    got = _run(TARA, tmp_path / "tara", n, steps)
    for c, name in enumerate(COLUMNS[:-1]):
        scale = np.max(np.abs(want[:, c]))
        err = np.max(np.abs(got[:, c] - want[:, c])) / scale
        print(f"  {name}: max rel diff {err:.2e}", flush=True)
        assert err < 1e-10, (name, err)
    # div B is roundoff in both (the initial B is uniform and the induction
    # equation keeps it solenoidal): below 1e-12 of B^2's scale in both
    bscale = np.max(np.abs(want[:, 4]))
    assert np.max(np.abs(got[:, -1])) < 1e-12 * bscale and np.max(np.abs(want[:, -1])) < 1e-12 * bscale


@pytest.mark.parametrize("n,steps", [(16, 20)])
def test_device_diagnostics_vs_reference_program(tmp_path, n, steps):
    """The device-resident path with the device diagnostics: the reference's
    ABC-flow initial state (tara_mhd_ic_abc with main3com's parameters,
    src/main3com.f95:228-324), AB2 steps with the reference's 1/450 diffusion
    and CS = 1/M = 10, and the ten Energy.dat averages after every step
    (tara_mhd_record_energy) against the reference program's Energy.dat."""
    import tara_amd as T

    print(f"reference program {n}^3 x {steps} (CPU)", flush=True)
    want = _run(TARA_REF, tmp_path / "ref", n, steps)
    L = 2.0 * np.pi
    s = T.MhdSolver(n, n, n, T.MhdParams(L, L, L, 1.0 / 450.0, 1.0 / 450.0, 10.0, 1e-3))
    s.ic_abc(rho0=1.0, u0=0.1, b0=0.1, p0=1.0)
    print(f"device stepping + diagnostics {n}^3 x {steps}", flush=True)
    got = s.record_energy(steps, 1)
    s.close()
    for c, name in enumerate(COLUMNS[1:-1]):
        scale = np.max(np.abs(want[:, c + 1]))
        err = np.max(np.abs(got[:, c] - want[:, c + 1])) / scale
        print(f"  {name}: max rel diff {err:.2e}", flush=True)
        assert err < 1e-10, (name, err)
    bscale = np.max(np.abs(want[:, 4]))
    assert np.max(np.abs(got[:, -1])) < 1e-12 * bscale
