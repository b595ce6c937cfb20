"""The Fortran boundary: the reference's routine names (derive2csh, ab2csh,
derive2coh, ab2coh, derive2com, ab2com, derive3com, ab3com, ...) resolved to
tara_amd/fortran/*_dropin.f90 -> C ABI -> HIP, driven by Fortran time loops
(tests/fortran/drive_*.f90, the loops of the reference's main programs, with
their COMMON /comm/ layouts), against the reference build's trajectories."""
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


REFERENCE_NAMES = ("derive2csh", "ab2csh", "derive2css", "ab2css", "rk42css", "derive2coh", "ab2coh", "euler2coh",
                   "rk4_nk2coh", "rk4_ukx2coh", "rk4_uky2coh", "pc_nk2coh", "pc_ukx2coh", "pc_uky2coh",
                   "derive2com", "ab2com", "derive3com", "ab3com")


def test_fortran_driver_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "tara_amd", "libtara_fortran.so")],
                         capture_output=True, text=True).stdout
    for sym in REFERENCE_NAMES:
        assert f" T {sym}_" in out, sym


def _write_fields(path, dims, fields):
    with open(path, "wb") as f:
        np.array(list(dims) + [len(fields)], dtype="<i4").tofile(f)
        for a in fields:
            np.asfortranarray(a).T.astype("<c16").tofile(f)


def _read_fields(path, ndim):
    with open(path, "rb") as f:
        dims = np.fromfile(f, dtype="<i4", count=ndim)
        nf = int(np.fromfile(f, dtype="<i4", count=1)[0])
        shape = (dims[0] // 2 + 1,) + tuple(int(d) for d in dims[1:])
        n = int(np.prod(shape))
        return [np.asfortranarray(np.fromfile(f, dtype="<c16", count=n).reshape(shape[::-1]).T) for _ in range(nf)]


def test_fields_io_roundtrip(tmp_path):
    a = [np.asfortranarray(np.arange(9 * 16 * 4).reshape(9, 16, 4) * (1 + 2j) + k) for k in range(3)]
    _write_fields(tmp_path / "a.bin", (16, 16, 4), a)
    b = _read_fields(tmp_path / "a.bin", 3)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


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


def _driver(name):
    exe = os.path.join(ROOT, "tests", "fortran", name)
    assert os.path.exists(exe), f"tests/fortran/{name} not built (run __graft_entry__.build())"
    return exe


@pytest.mark.gpu
def test_fortran_dropin_coh(golden_fam2d, tmp_path):
    n = 32
    _write_fields(tmp_path / "in.bin", (n, n), [golden_fam2d[f"coh_{n}_in{k}"] for k in range(3)])
    print("fortran coh", flush=True)
    subprocess.run([_driver("drive_coh"), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "200", "1.0d-4",
                    "2.0d-4", "3.77d0"], check=True, timeout=120)
    got = _read_fields(tmp_path / "out.bin", 2)
    for k in range(3):
        assert rel_l2(got[k], golden_fam2d[f"cohtraj_{n}_w{k}"]) < 1e-10


@pytest.mark.gpu
def test_fortran_dropin_com(golden_fam2d, tmp_path):
    n = 32
    _write_fields(tmp_path / "in.bin", (n, n), [golden_fam2d[f"com_{n}_in{k}"] for k in range(6)])
    print("fortran com", flush=True)
    subprocess.run([_driver("drive_com"), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "200", "1.0d-4",
                    "1.0d-3", "1.0d-4", "2.0d0"], check=True, timeout=120)
    got = _read_fields(tmp_path / "out.bin", 2)
    for k in range(6):
        assert rel_l2(got[k], golden_fam2d[f"comtraj_{n}_w{k}"]) < 1e-10


@pytest.mark.gpu
def test_fortran_dropin_mhd_one_step(golden, tmp_path):
    """One derive3com + ab3com step through the Fortran names with the
    reference's 1/450 diffusion: from zero history the update is
    state + 3/2 dt f, f = the reference derive3com vector; E is unchanged."""
    n, dt = 16, 1e-3
    st = [golden[f"d3com_in{k}"] for k in range(8)]
    _write_fields(tmp_path / "in.bin", (n, n, n), st)
    print("fortran mhd", flush=True)
    subprocess.run([_driver("drive_mhd"), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "1", "1.0d-3",
                    "10.0d0"], check=True, timeout=120)
    got = _read_fields(tmp_path / "out.bin", 3)
    assert np.array_equal(got[4], st[4])
    for k in (0, 1, 2, 3, 5, 6, 7):
        # the increment, so that the state's size does not hide an error in f
        assert rel_l2(got[k] - st[k], 1.5 * dt * golden[f"d3com_out{k}"]) < 1e-12, k


def test_dfftw_shim_exports_both_manglings():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "tara_amd", "libtara_dfftw.so")],
                         capture_output=True, text=True).stdout
    for base in ("dfftw_plan_dft_r2c_1d", "dfftw_plan_dft_c2r_1d", "dfftw_plan_dft_r2c_2d", "dfftw_plan_dft_c2r_2d",
                 "dfftw_plan_dft_r2c_3d", "dfftw_plan_dft_c2r_3d", "dfftw_execute", "dfftw_execute_dft_r2c",
                 "dfftw_execute_dft_c2r", "dfftw_destroy_plan", "dfftw_init_threads", "dfftw_plan_with_nthreads"):
        assert f" T {base}_\n" in out and f" T {base}__\n" in out, base


@pytest.mark.gpu
def test_dfftw_shim_vs_oracle(tmp_path):
    """FFTW's Fortran entry points, spelled as the reference spells them, served
    by libtara_dfftw.so: 2D, 3D and 1D r2c against the oracle, c2r round trips
    (unnormalised: N times the input)."""
    import oracle as O
    print("dfftw shim", flush=True)
    subprocess.run([_driver("drive_dfftw"), str(tmp_path / "out.bin")], check=True, timeout=120)
    raw = np.fromfile(tmp_path / "out.bin", dtype="<f8")
    off = 0

    def take(shape, cplx):
        nonlocal off
        n = int(np.prod(shape)) * (2 if cplx else 1)
        a = raw[off:off + n]
        off += n
        if cplx:
            a = a.view("<c16")
        return np.asfortranarray(a.reshape(shape[::-1]).T)

    r2, c2, b2 = take((32, 16), False), take((17, 16), True), take((32, 16), False)
    r3, c3, b3 = take((16, 8, 8), False), take((9, 8, 8), True), take((16, 8, 8), False)
    r1, c1, b1 = take((32,), False), take((17,), True), take((32,), False)
    assert off == raw.size
    assert rel_l2(c2, O.r2c(r2)) < 1e-14
    assert rel_l2(b2, 512 * r2) < 1e-14
    assert rel_l2(c3, O.r2c(r3)) < 1e-14
    assert rel_l2(b3, 1024 * r3) < 1e-14
    assert rel_l2(c1, np.fft.rfft(r1)) < 1e-14
    assert rel_l2(b1, 32 * r1) < 1e-14


@pytest.mark.gpu
def test_fortran_resident_mode(tmp_path):
    """The device-resident API from Fortran (module tara_hip_mod,
    tests/fortran/drive_resident.f90): the config-1 run with tracer particles
    kept on the GPU for 1001 steps against the reference program's snapshots
    (tests/golden/golden_snap.npz; particles that never enter the last row),
    and 20 rows of the 3D Energy.dat averages against the same calls made
    through the Python binding (the same library: equal bits)."""
    import tara_amd as T

    g = np.load(os.path.join(ROOT, "tests", "golden", "golden_snap.npz"))
    start = g["s64_start"]
    npart = start.shape[0]
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([npart], dtype="<i4").tofile(f)
        start[:, 0].astype("<f8").tofile(f)
        start[:, 1].astype("<f8").tofile(f)
    print("fortran resident", flush=True)
    subprocess.run([_driver("drive_resident"), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True,
                   timeout=300)
    out = np.fromfile(tmp_path / "out.bin", dtype="<f8")
    n2, nrec = 64, 20
    xp, yp = out[:npart], out[npart:2 * npart]
    fields = out[2 * npart:2 * npart + 3 * n2 * n2].reshape(3, n2, n2)  # each (x, y) column-major: [y][x]
    rows = out[2 * npart + 3 * n2 * n2:].reshape(nrec, 10)
    ok = ~g["s64_last_row"]
    p = g["s64_particles_1000"]
    assert np.max(np.abs(xp[ok] - p[ok, 0])) < 1e-10 and np.max(np.abs(yp[ok] - p[ok, 1])) < 1e-10
    f = g["s64_fields_1000"]
    for c in range(3):
        got = fields[c].T  # (x, y)
        ref = f[:, 2 + c].reshape(n2, n2)
        assert np.max(np.abs(got - ref)) < 1e-10 * np.max(np.abs(ref)), c
    L = 2.0 * np.pi
    s = T.MhdSolver(16, 16, 16, T.MhdParams(L, L, L, 1.0 / 450.0, 1.0 / 450.0, 10.0, 1e-3))
    s.ic_abc(1.0, 0.1, 0.1, 1.0, 1.0, 1.0, 1.0, 1.0)
    want = s.record_energy(nrec, 1)
    s.close()
    assert np.array_equal(rows, np.asarray(want).reshape(nrec, 10))
