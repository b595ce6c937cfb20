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

The config-1 program (dispatcher -> tara2dHydSer, src/main2csh.f95) runs
the same way: tests/fortran/_mains/tara_snap, the main compiled with its two
snapshot OPEN statements made flang-safe (flang's runtime stops on their
re-opening of a connected unit with STATUS='unknown', src/main2csh.f95:177,
:268; gfortran tolerates it), derive2csh / ab2csh through the wrappers and its
three c2r per step through the dfftw shim; its text snapshots (fields and
tracer particles) are compared with those of the program built against the
reference routines (tests/golden/golden_snap.npz).

All binaries are built by __graft_entry__.build() where the reference checkout
is present.  The reference allocates its AB2 history uninitialised
(src/main3com.f95 allocates d_*_dt_old, src/main2csh.f95 dt_omegak_old
without setting them); at these sizes that memory is fresh from the kernel,
hence zero, the zero history both paths assume (MALLOC_MMAP_THRESHOLD_ is
also set, where glibc honours it)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

MAINS = os.path.join(ROOT, "tests", "fortran", "_mains")
TARA, TARA_REF = os.path.join(MAINS, "tara"), os.path.join(MAINS, "tara_ref")
TARA_SNAP = os.path.join(MAINS, "tara_snap")
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


def _unlimited_stack():
    """derive3com keeps ~130 automatic arrays on the stack (29 MB at 24^3)."""
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def _run(exe, where, n, steps):
    where.mkdir()
    (where / "output").mkdir()
    (where / "input.ini").write_text(INI.format(n=n, tmax=repr((steps - 0.5) * 1e-3)))
    env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="4096")
    r = subprocess.run([exe, "input.ini"], cwd=where, env=env, capture_output=True, text=True, timeout=600,
                       preexec_fn=_unlimited_stack)
    assert r.returncode == 0, (exe, r.stdout[-2000:], r.stderr[-2000:])
    text = (where / "output" / "Energy.dat").read_text().replace("D", "E")
    vals = np.array([float(v) for v in text.split()])
    assert vals.size == steps * len(COLUMNS), (exe, vals.size)
    return vals.reshape(steps, len(COLUMNS))


@pytest.mark.parametrize("n,steps", [(16, 20), (24, 6)])
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


def _read_fort(path, ncol):
    with open(path) as f:
        return np.array(f.read().split(), dtype=float).reshape(-1, ncol)


@pytest.mark.skipif(not os.path.exists(TARA_SNAP), reason="config-1 program not linked")
def test_reference_main2csh_through_tara_vs_reference_program(tmp_path):
    """The config-1 program linked against tara_amd (GPU) on a 64^2 shear
    layer with 256 tracer particles, 1001 loop trips, against the same
    program built against the reference routines and FFTW: its snapshots of
    x, y, omega, ux, uy at t = 0 and t = 1000 to 1e-10 of each field's scale,
    and the particle positions (those that never enter the last row, where
    the program reads past its arrays into the heap) to 1e-10."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "golden_snap.npz"))
    (tmp_path / "output").mkdir()
    (tmp_path / "input.ini").write_text(str(g["s64_ini"]))
    print("main2csh 64^2 x 1001 linked against tara_amd (GPU)", flush=True)
    env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="4096")
    r = subprocess.run([TARA_SNAP, "input.ini"], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    ok = ~g["s64_last_row"]
    for t in (0, 1000):
        f = _read_fort(tmp_path / "output" / f"fort.{t + 100:08d}", 5)
        want = g[f"s64_fields_{t}"]
        assert np.array_equal(f[:, :2], want[:, :2])
        for c in (2, 3, 4):
            err = np.max(np.abs(f[:, c] - want[:, c])) / np.max(np.abs(want[:, c]))
            print(f"  t={t} column {c}: max rel diff {err:.2e}", flush=True)
            assert err < 1e-10, (t, c, err)
        p = _read_fort(tmp_path / "output" / f"fort.{t + 105:08d}", 2)
        want = g[f"s64_particles_{t}"]
        err = np.max(np.abs(p[ok] - want[ok]))
        print(f"  t={t} particles: max diff {err:.2e}", flush=True)
        assert err < 1e-10, (t, err)


@pytest.mark.skipif(not (os.path.exists(TARA_SNAP) and os.path.exists(TARA_SNAP + "_ref")),
                    reason="config-1 programs not linked")
def test_reference_main2csh_any_grid_through_tara(tmp_path):
    """The config-1 program on a 96 x 80 grid (not powers of two: derive2csh
    and the main's own c2r through the mixed-radix path), one loop trip,
    against the program built against the reference routines and FFTW: the
    t = 0 snapshot (fields, and the particles that do not start in the last
    row, where the program reads past its arrays into the heap)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_golden as G

    ini = G.snap_ini(96, 256, 1).replace("Ny = 96", "Ny = 80")
    out = {}
    for exe, name in ((TARA_SNAP + "_ref", "ref"), (TARA_SNAP, "tara")):
        d = tmp_path / name
        (d / "output").mkdir(parents=True)
        (d / "input.ini").write_text(ini)
        print(f"main2csh 96x80 x 1: {name}", flush=True)
        r = subprocess.run([exe, "input.ini"], cwd=d, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (name, r.stdout[-2000:], r.stderr[-2000:])
        out[name] = (_read_fort(d / "output" / "fort.00000100", 5), _read_fort(d / "output" / "fort.00000105", 2))
    (fr, pr), (ft, pt) = out["ref"], out["tara"]
    assert fr.shape == (96 * 80, 5) and np.array_equal(ft[:, :2], fr[:, :2])
    for c in (2, 3, 4):
        err = np.max(np.abs(ft[:, c] - fr[:, c])) / np.max(np.abs(fr[:, c]))
        print(f"  column {c}: max rel diff {err:.2e}", flush=True)
        assert err < 1e-10, (c, err)
    start = G.rand_starts(256, 2 * np.pi)
    ok = np.trunc(start[:, 1] / (2 * np.pi / 80)) < 79
    assert np.max(np.abs(pt[ok] - pr[ok])) < 1e-10
