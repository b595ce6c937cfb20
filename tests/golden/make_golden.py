"""Writes the golden fixtures under tests/golden/ from the REFERENCE build
(oracle/_ref/libtara_ref.so: TARA's own Fortran routines + the vendored FFTW
3.3.9, compiled from /root/reference by `make -C oracle ref`).

The reference has no tests, fixtures or golden files of its own; these vectors
are produced by running its routines on the inputs below.  Run here (where
/root/reference exists); the .npz files travel with the repository, the
reference does not.

    ulimit -s unlimited && python tests/golden/make_golden.py [core|fam2d|long|snap|all]

(derive3com keeps ~130 automatic arrays on the stack, hence the stack limit.)
"""
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

PI = 3.14159265358979323846


def smooth_field(rng, n, nmodes=6, kmax=3):
    x = np.arange(n) * 2 * PI / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.zeros((n, n, n))
    for _ in range(nmodes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        f += rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 2 * PI))
    return f


def mhd_state(seed, n):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.1 * smooth_field(rng, n) / 6
    u = [0.1 * smooth_field(rng, n) for _ in range(3)]
    B = [0.1 * smooth_field(rng, n) for _ in range(3)]
    E = np.ones((n, n, n))
    reals = [rho, rho * u[0], rho * u[1], rho * u[2], E, B[0], B[1], B[2]]
    return [O.r2c(np.asfortranarray(r)) for r in reals]


def random_spec(seed, n):
    rng = np.random.default_rng(seed)
    return O.r2c(np.asfortranarray(rng.standard_normal((n, n))))


def coh_state(seed, n):
    """n = 1 + small perturbation, smooth ux, uy (Lx = Ly = 2 pi)."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) * 2 * PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")

    def smooth():
        f = np.zeros((n, n))
        for _ in range(5):
            k = rng.integers(-3, 4, size=2)
            f += rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + rng.uniform(0, 2 * PI))
        return f

    dens = 1.0 + 0.05 * smooth()
    return X, Y, dens, 0.2 * smooth(), 0.2 * smooth(), smooth


def fam2d(out):
    """2D compressible families from the reference build: derive2coh, derive2com
    (forcing off: A = kf = 0, its x, y arrays being uninitialised), the coh and com
    AB2 loops (src/main2coh.f95:278-300, src/main2com.f95:364-380; history zero),
    and the pointwise steppers."""
    pi, L = PI, 2 * PI
    coh = dict(pi=pi, lx=L, ly=L, mux=2e-4, muy=2e-4, cs=3.77)
    com = dict(pi=pi, lx=L, ly=L, mu=1e-3, eta=1e-4, cs=2.0)
    for n in (32, 64):
        _, _, dens, ux, uy, smooth = coh_state(200 + n, n)
        st = [O.r2c(np.asfortranarray(a)) for a in (dens, ux, uy)]
        f = O.ref_derive2coh(st, n, **coh)
        for k in range(3):
            out[f"coh_{n}_in{k}"] = st[k]
            out[f"coh_{n}_out{k}"] = f[k]
        rho = dens
        E = 2.0 + 0.1 * smooth()
        bx, by = 0.1 * smooth(), 0.1 * smooth()
        st = [O.r2c(np.asfortranarray(a)) for a in (rho, rho * ux, rho * uy, E, bx, by)]
        f = O.ref_derive2com(st, n, **com)
        assert all(np.isfinite(x).all() for x in f)
        for k in range(6):
            out[f"com_{n}_in{k}"] = st[k]
            out[f"com_{n}_out{k}"] = f[k]
    # 200-step AB2 loops at 32^2
    n, dt = 32, 1e-4
    st = [out[f"coh_{n}_in{k}"] for k in range(3)]
    hist = [np.zeros_like(st[0], order="F") for _ in range(3)]
    for _ in range(200):
        f = O.ref_derive2coh(st, n, **coh)
        st = O.ref_ab2coh(st, hist, f, n, dt)
        hist = f
    for k in range(3):
        out[f"cohtraj_{n}_w{k}"] = st[k]
    st = [out[f"com_{n}_in{k}"] for k in range(6)]
    hist = [np.zeros_like(st[0], order="F") for _ in range(6)]
    for _ in range(200):
        f = O.ref_derive2com(st, n, **com)
        st = O.ref_ab2com(st, hist, f, n, dt)
        hist = f
    for k in range(6):
        out[f"comtraj_{n}_w{k}"] = st[k]
    # pointwise steppers at 32^2
    st = [out[f"coh_{n}_in{k}"] for k in range(3)]
    sl = [random_spec(300 + k, n) for k in range(3)]
    fo = [random_spec(310 + k, n) for k in range(3)]
    for k in range(3):
        out[f"pt_slope{k}"] = sl[k]
        out[f"pt_fold{k}"] = fo[k]
    for k, a in enumerate(O.ref_ab2coh(st, fo, sl, n, 1e-3)):
        out[f"pt_ab2coh{k}"] = a
    for k, a in enumerate(O.ref_euler2coh(st, sl, n, 1e-3)):
        out[f"pt_euler{k}"] = a
    out["pt_rk4nk"] = O.ref_nk_stepper("rk4_nk2coh_", st, sl, n, 1e-3, **coh)
    out["pt_pcnk"] = O.ref_nk_stepper("pc_nk2coh_", st, sl, n, 1e-3, **coh)


def main_fam2d():
    out = {}
    fam2d(out)
    np.savez_compressed(os.path.join(HERE, "golden_fam2d.npz"), **out)
    print("wrote", len(out), "arrays to golden_fam2d.npz")


def main():
    meta = dict(
        generator="tests/golden/make_golden.py",
        reference="RupakMukherjee/TARA src/*.f95 + lib/fftw 3.3.9 (oracle/Makefile)",
        fc=subprocess.run(["amdflang", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
        cc=subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
        host=platform.machine(),
    )
    out = {}
    # Config-1 trajectory (src/main2csh.f95 loop), 64^2, nu = 0.01, dt = 1e-4,
    # shear layer W0 = 25, m0 = 2; d as-run (3.0 after list-directed parsing of
    # "3.0d0/128.0d0") and the intended 3/128.  History starts at zero.
    for tag, d in (("asrun", 3.0), ("intent", 3.0 / 128.0)):
        omega, wk = O.ref_ic2csh(64, 64, 2.0, 2.0, 25.0, 2.0, d)
        out[f"c1_{tag}_omega0"] = omega
        out[f"c1_{tag}_w0"] = wk
        w, h = wk, O.zeros_spec(64, 64)
        done = 0
        for target in (1, 1000, 2000):
            w, h = O.ref_ab2csh_steps(w, h, 64, 0.01, 1e-4, target - done)
            done = target
            out[f"c1_{tag}_w{target}"] = w
        out[f"c1_{tag}_hist2000"] = h
    # RK4 over derive2csh (harness), 64^2 shear layer, 200 steps.
    out["rk4_w0"] = out["c1_asrun_w0"]
    out["rk4_w200"] = O.ref_rk4csh_steps(out["rk4_w0"], 64, 0.01, 1e-4, 200)
    # Single derive2csh calls.
    for n in (64, 128):
        _, ic = O.ref_ic2csh(n, n, 2.0, 2.0, 25.0, 2.0, 3.0 / 128.0)
        rnd = random_spec(100 + n, n)
        out[f"d2csh_{n}_ic_in"] = ic
        out[f"d2csh_{n}_ic_out"] = O.ref_derive2csh(ic, n, 0.01)
        out[f"d2csh_{n}_rnd_in"] = rnd
        out[f"d2csh_{n}_rnd_out"] = O.ref_derive2csh(rnd, n, 1e-3)
    # Single ab2csh call.
    a, b, c = random_spec(1, 64), random_spec(2, 64), random_spec(3, 64)
    out["ab2_in_w"], out["ab2_in_fo"], out["ab2_in_fn"] = a, b, c
    out["ab2_out"] = O.ref_ab2csh(a, b, c, 64, 1e-3)
    # Single derive3com call at 16^3 (diffusion 1/450 as hard-wired, CS = 10).
    st = mhd_state(7, 16)
    f = O.ref_derive3com(st, 16, 2 * PI, 2 * PI, 2 * PI, cs=10.0)
    for k in range(8):
        out[f"d3com_in{k}"] = st[k]
        if k != 4:
            out[f"d3com_out{k}"] = f[k]
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    with open(os.path.join(HERE, "golden_meta.txt"), "w") as fh:
        for k, v in meta.items():
            fh.write(f"{k}: {v}\n")
    print("wrote", len(out), "arrays")


def block(w, m):
    """The modes kx <= m, |ky| <= m of a (Nh, N) half spectrum, ky ascending
    from -m to m: shape (m + 1, 2 m + 1)."""
    n = w.shape[1]
    return np.asfortranarray(np.concatenate([w[:m + 1, n - m:], w[:m + 1, :m + 1]], axis=1))


def main_long():
    """Long-horizon and 3D fixtures (golden_long.npz):
    * ab3com: one call of the reference ab3com (src/ab3com.f95:6-66) with
      non-zero history at 8^3;
    * d3mueta_16: the patched reference derive3com_mueta (diffusion mu on the
      momentum, eta on the induction, oracle/Makefile) at mu = 0.02,
      eta = 0.005 on the 16^3 state of golden.npz (d3com_in*);
    * c2_512 / c2_1024: the RK4 harness over the reference derive2csh for
      1000 steps at 512^2 and 1024^2, nu = 1e-4, dt = 2e-4, from the Philox
      random-phase IC (oracle.ic_random, seed 1, k0 = 4, unit energy): the
      final field (512^2) and, at 1024^2, the block kx <= 128, |ky| <= 128
      plus the full field's L2 norm and the norm outside the block;
    * mhd_32: the reference 3D AB2 loop (src/main3com.f95:536-617) over
      derive3com_mueta + ab3com for 500 steps at 32^3, nu = eta = 1e-3,
      dt = 5e-4, CS = 10, from oracle.ic_random3d(32, seed 1, k0 3, energies
      1.0 and 0.5): the 7 live fields at step 500."""
    out = {}
    n = 8
    rng = np.random.default_rng(11)

    def rspec():
        return O.r2c(np.asfortranarray(rng.standard_normal((n, n, n))))

    st, fo, fn = [rspec() for _ in range(8)], [rspec() for _ in range(8)], [rspec() for _ in range(8)]
    res = O.ref_ab3com(st, fo, fn, n, 1e-3)
    for k in range(8):
        out[f"ab3_st{k}"], out[f"ab3_fo{k}"], out[f"ab3_fn{k}"], out[f"ab3_out{k}"] = st[k], fo[k], fn[k], res[k]
    st = mhd_state(7, 16)
    f = O.ref_derive3com(st, 16, 2 * PI, 2 * PI, 2 * PI, cs=10.0, mu=0.02, eta=0.005)
    for k in (0, 1, 2, 3, 5, 6, 7):
        out[f"d3mueta_16_out{k}"] = f[k]
    L = 2 * PI
    ic = O.ic_random3d(32, seed=1, k0=3.0, e_kin=1.0, e_mag=0.5)
    fin, _ = O.ref_mhd_ab2_steps(ic, 32, L, L, L, 1e-3, 1e-3, 5e-4, 500, cs=10.0)
    for k in (0, 1, 2, 3, 5, 6, 7):
        out[f"mhd_32_w500_{k}"] = fin[k]
    print("3D done", flush=True)
    for nn in (512, 1024):
        w0 = O.ic_random(nn, nn, 1, 4.0, 1.0)
        w = O.ref_rk4csh_steps(w0, nn, 1e-4, 2e-4, 1000)
        if nn == 512:
            out["c2_512_w1000"] = w
        else:
            b = block(w, 128)
            out["c2_1024_block128"] = b
            out["c2_1024_norm"] = np.array(np.linalg.norm(w))
            inside = np.zeros(w.shape, bool)
            inside[:129, :129] = True
            inside[:129, nn - 128:] = True
            out["c2_1024_norm_outside"] = np.array(np.linalg.norm(w[~inside]))
        print(nn, "done", flush=True)
    np.savez_compressed(os.path.join(HERE, "golden_long.npz"), **out)
    print("wrote", len(out), "arrays to golden_long.npz")


SNAP_INI = """[dimension]
dim = 2
[architecture]
arch = cpu
[parallel]
mode = serial
[type]
domain = hydro
[problem]
hydrotype = bare
[particle]
Np = 256
[grid]
Nx = 64
Ny = 64
Nz = 64
[length]
Lx = 2.0d0
Ly = 2.0d0
Lz = 2.0d0
[resistivity]
nu = 0.010d0
eta = 0.0d0
[time]
time_min = 0.0d0
time_max = 1.0005d0
dt = 0.0010d0
[initialProfile]
W0 = 25.0d0
m0 = 2.0d0
d = 0.0234375d0
[initialB]
B0 = 0.0d0
"""


def read_fort(path, ncol):
    """A list-directed text file (records may wrap) as an (n, ncol) array."""
    with open(path) as f:
        return np.array(f.read().split(), dtype=float).reshape(-1, ncol)


def snap_ini(n, np_, trips):
    """input.ini of the config-1 program (dispatcher -> tara2dHydSer): n^2,
    np_ particles, `trips` loop trips (time_max between two multiples of dt)."""
    return (SNAP_INI.replace("Nx = 64", f"Nx = {n}").replace("Ny = 64", f"Ny = {n}")
            .replace("Np = 256", f"Np = {np_}").replace("time_max = 1.0005d0", f"time_max = {(trips - 0.5) * 1e-3!r}d0"))


def run_snap(exe, ini, times):
    """Runs a build of the config-1 program on `ini` in a scratch directory;
    returns {t: (fields, particles)} of its text snapshots at the given t."""
    import shutil
    import tempfile
    work = tempfile.mkdtemp()
    try:
        os.mkdir(os.path.join(work, "output"))
        with open(os.path.join(work, "input.ini"), "w") as f:
            f.write(ini)
        subprocess.run([exe, "input.ini"], cwd=work, check=True, capture_output=True, timeout=600)
        return {t: (read_fort(os.path.join(work, "output", f"fort.{t + 100:08d}"), 5),
                    read_fort(os.path.join(work, "output", f"fort.{t + 105:08d}"), 2)) for t in times}
    finally:
        shutil.rmtree(work)


def rand_starts(np_, lx):
    """main2csh's start positions (src/main2csh.f95:93-96): xp = RAND() Lx,
    yp = RAND() Ly in turn, from RAND of libtara_gnucompat, the generator the
    linked program calls, at its initial state."""
    import ctypes
    root = os.path.dirname(os.path.dirname(HERE))
    g = ctypes.CDLL(os.path.join(root, "tara_amd", "libtara_gnucompat.so"))
    g.rand_.restype = ctypes.c_float
    g.srand_(ctypes.byref(ctypes.c_int(1)))
    start = np.zeros((np_, 2))
    for i in range(np_):
        start[i, 0] = float(g.rand_()) * lx
        start[i, 1] = float(g.rand_()) * lx
    return start


def main_snap():
    """The reference's config-1 program itself (src/main.f95 -> tara2dHydSer,
    src/main2csh.f95), built against the reference routines and FFTW 3.3.9
    (tests/fortran/_mains/tara_snap_ref; its snapshot OPENs made flang-safe,
    see tests/fortran/Makefile.mains), and its text snapshots (fields: x, y,
    omega, ux, uy per point; particles: x, y):
      s64_*  : 64^2 shear layer, 256 particles, 1001 loop trips, snapshots at
               t = 0 and t = 1000;
      s128_* : 128^2, 2048 particles, one trip, the particles only.  At this
               size the program's arrays are 128 KiB, which glibc maps with
               mmap, so the tracer code's reads past their end (last row) read
               the zero-filled rest of the last page; at 64^2 they read
               whatever the heap holds next (the following array).
    The AB2 history the program leaves uninitialised is zero in both (memory
    fresh from the kernel)."""
    root = os.path.dirname(os.path.dirname(HERE))
    exe = os.path.join(root, "tests", "fortran", "_mains", "tara_snap_ref")
    subprocess.run(["make", "-s", "-f", os.path.join(root, "tests", "fortran", "Makefile.mains")], check=True)
    lx = 2.0 * PI
    out = {}
    ini = snap_ini(64, 256, 1001)
    r = run_snap(exe, ini, (0, 1000))
    out["s64_ini"] = np.array(ini)
    out["s64_start"] = rand_starts(256, lx)
    for t in (0, 1000):
        out[f"s64_fields_{t}"], out[f"s64_particles_{t}"] = r[t]
    # which particles enter the last row during the run (there the program
    # reads the heap past its arrays): the oracle's trajectory, which equals
    # the program's for the others
    _, w = O.ic2csh(64, 64, d_in=0.0234375)
    hist = np.zeros_like(w, order="F")
    xp, yp = out["s64_start"][:, 0].copy(), out["s64_start"][:, 1].copy()
    last_row = np.zeros(256, bool)
    for _ in range(1001):
        w, hist = O.ab2csh_steps(w, hist, 64, 0.01, 1e-3, 1)
        ux, uy = O.csh_velocity(w, 64, lx, lx)
        last_row |= np.trunc(yp / (lx / 64)) >= 63
        xp, yp = O.tracer_step(ux, uy, xp, yp, lx, lx, 1e-3)
    out["s64_last_row"] = last_row
    ini = snap_ini(128, 2048, 1)
    r = run_snap(exe, ini, (0,))
    out["s128_ini"] = np.array(ini)
    out["s128_start"] = rand_starts(2048, lx)
    out["s128_particles_0"] = r[0][1]
    np.savez_compressed(os.path.join(HERE, "golden_snap.npz"), **out)
    print("wrote", len(out), "arrays to golden_snap.npz")


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("snap", "all"):
        main_snap()
    if what in ("long", "all"):
        main_long()
    if what in ("core", "all"):
        main()
    if what in ("fam2d", "all"):
        main_fam2d()
