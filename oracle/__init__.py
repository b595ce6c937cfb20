"""TEST INFRASTRUCTURE ONLY -- the CPU oracle.

ctypes bindings for
  * ``liboracle.so``        : the plain-C restatement (tara_oracle.c) of the
                              reference arithmetic, and
  * ``_ref/libtara_ref.so`` : the reference's own Fortran routines linked with
                              the vendored FFTW 3.3.9, built from the reference
                              sources by ``oracle/Makefile`` (``make ref``).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, and only as the checker / CPU baseline.  The product path
(``tara_amd``) never imports it.

Arrays follow the reference layout: Fortran column-major complex128 half
spectra of shape (Nh, Ny[, Nz]) with Nh = Nx//2 + 1 (numpy ``order='F'``), real
fields (Nx, Ny[, Nz]) likewise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

_orc = None
_ref = None


def _ptr(a):
    assert a.flags.f_contiguous or a.ndim == 1, "arrays must be Fortran-ordered"
    return a.ctypes.data_as(_dp)


def lib():
    """Load liboracle.so (the C restatement)."""
    global _orc
    if _orc is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        _orc = C.CDLL(path)
        _orc.orc_dft.argtypes = [C.c_int, _dp, C.c_long, C.c_int]
        for n in ("orc_r2c_2d", "orc_c2r_2d"):
            getattr(_orc, n).argtypes = [C.c_int, C.c_int, _dp, _dp]
        for n in ("orc_r2c_3d", "orc_c2r_3d"):
            getattr(_orc, n).argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        _orc.orc_derive2csh.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp]
        _orc.orc_ab2csh.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        _orc.orc_ab2csh_steps.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        _orc.orc_rk4csh_steps.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp]
        _orc.orc_derive2css.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        _orc.orc_rk4css_steps.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4 + [C.c_int, _dp]
        _orc.orc_ic2csh.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [_dp, _dp]
        _orc.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
        _orc.orc_ic_random.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_double, C.c_double, _dp]
        _orc.orc_derive2coh.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _orc.orc_derive2com.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _orc.orc_derive3com.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _orc.orc_ab3com.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double] + [C.c_void_p] * 4
        _orc.orc_ic_random3d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint] + [C.c_double] * 3 + [C.c_void_p]
    return _orc


def ref_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libtara_ref.so"))


def ref():
    """Load _ref/libtara_ref.so (reference Fortran + vendored FFTW)."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libtara_ref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the reference is present")
        _ref = C.CDLL(path)
        _ref.ref_ic2csh.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [_dp, _dp]
        _ref.ref_ab2csh_steps.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        _ref.ref_rk4csh_steps.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp]
        _ref.ref_mhd_ab2_steps.argtypes = [C.c_int] * 5 + [_dp, _dp]
    return _ref


def zeros_spec(nx, ny, nz=None):
    shape = (nx // 2 + 1, ny) if nz is None else (nx // 2 + 1, ny, nz)
    return np.zeros(shape, dtype=np.complex128, order="F")


# --------------------------------------------------------------- transforms
def r2c(real):
    real = np.asfortranarray(real, dtype=np.float64)
    if real.ndim == 2:
        nx, ny = real.shape
        out = zeros_spec(nx, ny)
        lib().orc_r2c_2d(nx, ny, _ptr(real), _ptr(out))
    else:
        nx, ny, nz = real.shape
        out = zeros_spec(nx, ny, nz)
        lib().orc_r2c_3d(nx, ny, nz, _ptr(real), _ptr(out))
    return out


def c2r(spec, nx):
    spec = np.asfortranarray(spec, dtype=np.complex128)
    if spec.ndim == 2:
        ny = spec.shape[1]
        out = np.zeros((nx, ny), order="F")
        lib().orc_c2r_2d(nx, ny, _ptr(spec), _ptr(out))
    else:
        ny, nz = spec.shape[1:]
        out = np.zeros((nx, ny, nz), order="F")
        lib().orc_c2r_3d(nx, ny, nz, _ptr(spec), _ptr(out))
    return out


def dft(z, sign):
    z = np.array(z, dtype=np.complex128)
    lib().orc_dft(z.size, z.ctypes.data_as(_dp), 1, sign)
    return z


# ----------------------------------------------------------------- 2D hydro
def derive2csh(w, nx, nu):
    w = np.asfortranarray(w, dtype=np.complex128)
    out = np.zeros_like(w, order="F")
    lib().orc_derive2csh(nx, w.shape[1], nu, _ptr(w), _ptr(out))
    return out


def ab2csh(w, f_old, f_new, nx, dt):
    w, f_old, f_new = (np.asfortranarray(a, dtype=np.complex128) for a in (w, f_old, f_new))
    out = np.zeros_like(w, order="F")
    lib().orc_ab2csh(nx, w.shape[1], dt, _ptr(w), _ptr(out), _ptr(f_old), _ptr(f_new))
    return out


def ab2csh_steps(w, hist, nx, nu, dt, nsteps):
    w = np.array(w, dtype=np.complex128, order="F")
    hist = np.array(hist, dtype=np.complex128, order="F")
    lib().orc_ab2csh_steps(nx, w.shape[1], nu, dt, nsteps, _ptr(w), _ptr(hist))
    return w, hist


def rk4csh_steps(w, nx, nu, dt, nsteps):
    w = np.array(w, dtype=np.complex128, order="F")
    lib().orc_rk4csh_steps(nx, w.shape[1], nu, dt, nsteps, _ptr(w))
    return w


def derive2css(w, nx, alpha, nu, lam):
    w = np.asfortranarray(w, dtype=np.complex128)
    out = np.zeros_like(w, order="F")
    lib().orc_derive2css(nx, w.shape[1], alpha, nu, lam, _ptr(w), _ptr(out))
    return out


def rk4css_steps(w, nx, alpha, nu, lam, dt, nsteps):
    w = np.array(w, dtype=np.complex128, order="F")
    lib().orc_rk4css_steps(nx, w.shape[1], alpha, nu, lam, dt, nsteps, _ptr(w))
    return w


def ic2csh(nx, ny, lx_in=2.0, ly_in=2.0, w0=25.0, m0=2.0, d_in=3.0):
    omega = np.zeros((nx, ny), order="F")
    wk = zeros_spec(nx, ny)
    lib().orc_ic2csh(nx, ny, lx_in, ly_in, w0, m0, d_in, _ptr(omega), _ptr(wk))
    return omega, wk


# --------------------------------------------------------- tracer particles
def csh_velocity(w, nx, lx, ly):
    """ux, uy of main2csh's per-step work (src/main2csh.f95:128-167): the
    driver's wavenumbers kx = 2 pi (i-1)/Lx, ky = 2 pi (j-1)/Ly for
    j <= Ny/2 else 2 pi ((j-1)-Ny)/Ly, psi = omega/(kx^2+ky^2) with
    psi(1,1) = 0, ukx = +i ky psi, uky = -i kx psi, unnormalised c2r
    (FFTW semantics, as c2r here) divided by Nx*Ny."""
    w = np.asfortranarray(w, dtype=np.complex128)
    nh, ny = w.shape
    kx = (2.0 * np.pi * np.arange(nh, dtype=float) / lx)[:, None]
    j = np.arange(ny)
    ky = (2.0 * np.pi * np.where(j < ny // 2, j, j - ny).astype(float) / ly)[None, :]
    k2 = kx * kx + ky * ky
    k2[0, 0] = 1.0
    psi = (w.real / k2) + 1j * (w.imag / k2)
    psi[0, 0] = 0.0
    ukx = (-(ky * psi.imag)) + 1j * (ky * psi.real)
    uky = (kx * psi.imag) + 1j * (-(kx * psi.real))
    n = float(nx) * float(ny)
    return c2r(ukx, nx) / n, c2r(uky, nx) / n


def tracer_step(ux, uy, xp, yp, lx, ly, dt, periodic=False):
    """One step of the tracer particles (src/main2csh.f95:187-265), returning
    the new (xp, yp).  Cell m = aint(xp/dx) (0-based here), bilinear
    interpolation in the driver's expression (h = 1/(dx dy), x(m+1) = Lx in
    the last column, y(n+1) = Ly in the last row), x <- x + v dt, then
    x <- x - int(x/Lx) Lx and + Lx if <= 0.  Corners: with periodic=False as
    the driver addresses its column-major arrays (flat index m + n Nx, +1 for
    the right neighbour, +Nx for the upper one), a read past the end giving 0
    (undefined in the language; what the driver reads from the zero-filled
    rest of the last page when its arrays are mmap-allocated); with
    periodic=True the periodic images.  A cell index at or past Nx (xp = Lx) is taken as the
    last cell."""
    ux = np.asfortranarray(ux)
    uy = np.asfortranarray(uy)
    nx, ny = ux.shape
    dx, dy = lx / nx, ly / ny
    h = 1.0 / (dx * dy)
    xp = np.asarray(xp, dtype=float)
    yp = np.asarray(yp, dtype=float)
    m = np.clip(np.trunc(xp / dx).astype(np.int64), 0, nx - 1)
    n = np.clip(np.trunc(yp / dy).astype(np.int64), 0, ny - 1)
    xm, ym = m * dx, n * dy
    xm1 = np.where(m + 1 < nx, (m + 1) * dx, lx)
    yn1 = np.where(n + 1 < ny, (n + 1) * dy, ly)
    a = m + n * nx
    if periodic:
        m1, n1 = (m + 1) % nx, (n + 1) % ny
        b, c, d = m1 + n * nx, m + n1 * nx, m1 + n1 * nx
    else:
        b, c, d = a + 1, a + nx, a + nx + 1
    fx = np.concatenate([ux.ravel(order="F"), np.zeros(nx + 1)])
    fy = np.concatenate([uy.ravel(order="F"), np.zeros(nx + 1)])

    def interp(f):
        return (f[a] * (xm1 - xp) * (yn1 - yp) * h + f[b] * (xp - xm) * (yn1 - yp) * h
                + f[c] * (xm1 - xp) * (yp - ym) * h + f[d] * (xp - xm) * (yp - ym) * h)

    x = xp + interp(fx) * dt
    y = yp + interp(fy) * dt
    x = x - np.trunc(x / lx) * lx
    x = np.where(x <= 0.0, x + lx, x)
    y = y - np.trunc(y / ly) * ly
    y = np.where(y <= 0.0, y + ly, y)
    return x, y


# ------------------------------------------------------------------ 3D MHD
class MhdParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("lx", "ly", "lz", "mu", "eta", "cs", "dt")]


def _ptr_array(arrs):
    return (C.c_void_p * 8)(*[a.ctypes.data for a in arrs])


def derive3com(state, nx, lx, ly, lz, mu, eta, cs=10.0):
    """state: list of 8 (Nh,Ny,Nz) spectra (rho, rho_ux, rho_uy, rho_uz, E, Bx, By, Bz)."""
    state = [np.asfortranarray(s, dtype=np.complex128) for s in state]
    ny, nz = state[0].shape[1:]
    out = [np.zeros_like(state[0], order="F") for _ in range(8)]
    p = MhdParams(lx, ly, lz, mu, eta, cs, 0.0)
    lib().orc_derive3com(nx, ny, nz, C.byref(p), _ptr_array(state), _ptr_array(out))
    return out


def ab3com(state, f_old, f_new, nx, dt):
    state = [np.asfortranarray(s, dtype=np.complex128) for s in state]
    f_old = [np.asfortranarray(s, dtype=np.complex128) for s in f_old]
    f_new = [np.asfortranarray(s, dtype=np.complex128) for s in f_new]
    ny, nz = state[0].shape[1:]
    out = [np.zeros_like(state[0], order="F") for _ in range(8)]
    lib().orc_ab3com(nx, ny, nz, dt, _ptr_array(state), _ptr_array(out), _ptr_array(f_old), _ptr_array(f_new))
    return out


def ic_random3d(n, seed=1, k0=3.0, e_kin=1.0, e_mag=0.5):
    """Eight (Nh, N, N) spectra of the random-phase 3D state (orc_ic_random3d)."""
    out = [zeros_spec(n, n, n) for _ in range(8)]
    lib().orc_ic_random3d(n, n, n, seed, k0, e_kin, e_mag, _ptr_array(out))
    return out


# ------------------------------------------------- reference (FFTW) routines
def _i(v):
    return C.byref(C.c_int(v))


def _d(v):
    return C.byref(C.c_double(v))


def ref_derive2csh(w, nx, nu):
    w = np.array(w, dtype=np.complex128, order="F")
    nh, ny = w.shape
    dum = np.zeros_like(w, order="F")
    out = np.zeros_like(w, order="F")
    ref().derive2csh_(_i(nx), _i(ny), _i(nh), _d(nu), _d(0.0), _ptr(w), _ptr(dum), _ptr(dum), _ptr(out))
    return out


def ref_ab2csh(w, f_old, f_new, nx, dt):
    w, f_old, f_new = (np.array(a, dtype=np.complex128, order="F") for a in (w, f_old, f_new))
    nh, ny = w.shape
    out = np.zeros_like(w, order="F")
    ref().ab2csh_(_i(nx), _i(ny), _i(nh), _d(0.0), _d(0.0), _d(dt), _ptr(w), _ptr(out), _ptr(f_old), _ptr(f_new))
    return out


def ref_ic2csh(nx, ny, lx_in=2.0, ly_in=2.0, w0=25.0, m0=2.0, d_in=3.0):
    omega = np.zeros((nx, ny), order="F")
    wk = zeros_spec(nx, ny)
    ref().ref_ic2csh(nx, ny, lx_in, ly_in, w0, m0, d_in, _ptr(omega), _ptr(wk))
    return omega, wk


def ref_ab2csh_steps(w, hist, nx, nu, dt, nsteps):
    w = np.array(w, dtype=np.complex128, order="F")
    hist = np.array(hist, dtype=np.complex128, order="F")
    ref().ref_ab2csh_steps(nx, w.shape[1], nu, dt, nsteps, _ptr(w), _ptr(hist))
    return w, hist


def ref_rk4csh_steps(w, nx, nu, dt, nsteps):
    w = np.array(w, dtype=np.complex128, order="F")
    ref().ref_rk4csh_steps(nx, w.shape[1], nu, dt, nsteps, _ptr(w))
    return w


def philox4x32_10(ctr, key):
    c = (C.c_uint * 4)(*ctr)
    k = (C.c_uint * 2)(*key)
    o = (C.c_uint * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def ic_random(nx, ny, seed, k0=4.0, energy=1.0):
    wk = zeros_spec(nx, ny)
    lib().orc_ic_random(nx, ny, seed, k0, energy, _ptr(wk))
    return wk


class CohParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("pi", "lx", "ly", "mux", "muy", "cs")]


class ComParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("pi", "lx", "ly", "mu", "eta", "cs", "kf", "a")]


def derive2coh(state, nx, pi, lx, ly, mux, muy, cs):
    state = [np.array(s, dtype=np.complex128, order="F") for s in state]
    out = [np.zeros_like(state[0], order="F") for _ in range(3)]
    prm = CohParams(pi, lx, ly, mux, muy, cs)
    lib().orc_derive2coh(nx, state[0].shape[1], C.byref(prm), _ptr_array(state), _ptr_array(out))
    return out


def derive2com(state, nx, pi, lx, ly, mu, eta, cs, kf, a):
    state = [np.array(s, dtype=np.complex128, order="F") for s in state]
    out = [np.zeros_like(state[0], order="F") for _ in range(6)]
    prm = ComParams(pi, lx, ly, mu, eta, cs, kf, a)
    lib().orc_derive2com(nx, state[0].shape[1], C.byref(prm), _ptr_array(state), _ptr_array(out))
    return out


class _CommCoh(C.Structure):
    """COMMON /comm/ as src/main2coh.f95:29 lays it out."""
    _fields_ = [(n, C.c_double) for n in ("Lx", "Ly", "mux", "muy", "ms", "CS", "dt")]


class _CommCom(C.Structure):
    """COMMON /comm/ as src/main2com.f95:43 lays it out."""
    _fields_ = [(n, C.c_double) for n in ("Lx", "Ly", "spheat", "mu", "ms", "CS", "mu_0", "eta", "dt", "kf", "A")]


def ref_derive2coh(state, nx, pi, lx, ly, mux, muy, cs):
    """Calls the reference derive2coh (COMMON written with the coh layout)."""
    state = [np.array(s, dtype=np.complex128, order="F") for s in state]
    nh, ny = state[0].shape
    cm = _CommCoh.in_dll(ref(), "comm_")
    cm.Lx, cm.Ly, cm.mux, cm.muy, cm.ms, cm.CS, cm.dt = lx, ly, mux, muy, 1.0, cs, 0.0
    old = [np.zeros_like(state[0], order="F") for _ in range(3)]
    new = [np.zeros_like(state[0], order="F") for _ in range(3)]
    ref().derive2coh_(_i(nx), _i(ny), _i(nh), _d(pi), _d(0.0), *[_ptr(s) for s in state],
                      *[_ptr(o) for o in old], *[_ptr(n) for n in new])
    return new


def ref_derive2com(state, nx, pi, lx, ly, mu, eta, cs, kf=0.0, a=0.0):
    """Calls the reference derive2com (COMMON written with the com layout).  Its
    forcing reads uninitialised x, y arrays; with A = kf = 0 the forcing terms
    are zero whatever those hold (unless they hold non-finite values, which the
    caller checks)."""
    state = [np.array(s, dtype=np.complex128, order="F") for s in state]
    nh, ny = state[0].shape
    cm = _CommCom.in_dll(ref(), "comm_")
    cm.Lx, cm.Ly, cm.spheat, cm.mu, cm.ms, cm.CS = lx, ly, 5.0 / 3.0, mu, 1.0, cs
    cm.mu_0, cm.eta, cm.dt, cm.kf, cm.A = 1.0, eta, 0.0, kf, a
    old = [np.zeros_like(state[0], order="F") for _ in range(6)]
    new = [np.zeros_like(state[0], order="F") for _ in range(6)]
    args = [_i(nx), _i(ny), _i(nh), _d(pi), _d(0.0)] + [_ptr(s) for s in state]
    for o, n in zip(old, new):
        args += [_ptr(o), _ptr(n)]
    ref().derive2com_(*args)
    return new


def _ref_pointwise(name, nx, args_before, arrays, *, dt, comm):
    """Calls a pointwise reference stepper after writing dt into COMMON."""
    comm.dt = dt
    getattr(ref(), name)(*args_before, *[_ptr(a) for a in arrays])


def ref_ab2coh(state, f_old, f_new, nx, dt):
    state, f_old, f_new = ([np.array(s, dtype=np.complex128, order="F") for s in x] for x in (state, f_old, f_new))
    nh, ny = state[0].shape
    new = [np.zeros_like(state[0], order="F") for _ in range(3)]
    cm = _CommCoh.in_dll(ref(), "comm_")
    _ref_pointwise("ab2coh_", nx, [_i(nx), _i(ny), _i(nh), _d(np.pi), _d(0.0)],
                   state + new + f_old + f_new, dt=dt, comm=cm)
    return new


def ref_euler2coh(state, slope, nx, dt):
    state, slope = ([np.array(s, dtype=np.complex128, order="F") for s in x] for x in (state, slope))
    nh, ny = state[0].shape
    new = [np.zeros_like(state[0], order="F") for _ in range(3)]
    cm = _CommCoh.in_dll(ref(), "comm_")
    _ref_pointwise("euler2coh_", nx, [_i(nx), _i(ny), _i(nh), _d(np.pi), _d(0.0)],
                   state + new + slope, dt=dt, comm=cm)
    return new


def ref_nk_stepper(name, state, slope, nx, dt, pi, lx, ly, mux, muy, cs):
    """rk4_nk2coh_ / pc_nk2coh_: new density from (nk, ukx, uky) and the slopes.
    Their internal derive2coh calls run on the given state (results unused)."""
    state, slope = ([np.array(s, dtype=np.complex128, order="F") for s in x] for x in (state, slope))
    nh, ny = state[0].shape
    cm = _CommCoh.in_dll(ref(), "comm_")
    cm.Lx, cm.Ly, cm.mux, cm.muy, cm.ms, cm.CS, cm.dt = lx, ly, mux, muy, 1.0, cs, dt
    nk_new = np.zeros_like(state[0], order="F")
    getattr(ref(), name)(_i(nx), _i(ny), _i(nh), _d(pi), _d(0.0), _ptr(state[0]), _ptr(nk_new), _ptr(state[1]),
                         _ptr(state[2]), *[_ptr(s) for s in slope])
    return nk_new


def ref_ab2com(state, f_old, f_new, nx, dt):
    state, f_old, f_new = ([np.array(s, dtype=np.complex128, order="F") for s in x] for x in (state, f_old, f_new))
    nh, ny = state[0].shape
    new = [np.zeros_like(state[0], order="F") for _ in range(6)]
    cm = _CommCom.in_dll(ref(), "comm_")
    cm.dt = dt
    args = [_i(nx), _i(ny), _i(nh), _d(np.pi), _d(0.0)] + [_ptr(s) for s in state] + [_ptr(s) for s in new]
    for o, n in zip(f_old, f_new):
        args += [_ptr(o), _ptr(n)]
    ref().ab2com_(*args)
    return new


class _Comm(C.Structure):
    """COMMON /comm/ of src/main3com.f95:56 / src/derive3com.f95:67."""
    _pack_ = 4
    _fields_ = [("iret", C.c_int), ("thread_num", C.c_int)] + [
        (n, C.c_double) for n in ("time_max", "Lx", "Ly", "Lz", "dx", "dy", "dz", "spheat",
                                  "mu", "ms", "CS", "mu_0", "eta", "dt", "kf")]


def ref_comm():
    return _Comm.in_dll(ref(), "comm_")


def _set_comm3(nx, ny, nz, lx, ly, lz, cs, threads, mu=None, eta=None, dt=0.0):
    cm = ref_comm()
    cm.thread_num = threads
    cm.Lx, cm.Ly, cm.Lz = lx, ly, lz
    cm.CS = cs
    cm.kf = 1.0
    cm.dx, cm.dy, cm.dz = lx / nx, ly / ny, lz / nz
    cm.dt = dt
    if mu is not None:
        cm.mu, cm.eta = mu, eta
    return cm


def ref_derive3com(state, nx, lx, ly, lz, cs=10.0, threads=1, mu=None, eta=None):
    """Calls the reference derive3com (diffusion hard-wired to 1/450, Q9) or,
    with mu and eta given, derive3com_mueta: the same routine with its 1/450
    replaced by COMMON's mu (momentum) and eta (induction), built by
    oracle/Makefile from a patched copy of src/derive3com.f95."""
    state = [np.array(s, dtype=np.complex128, order="F") for s in state]
    nh, ny, nz = state[0].shape
    _set_comm3(nx, ny, nz, lx, ly, lz, cs, threads, mu, eta)
    old = [np.zeros_like(state[0], order="F") for _ in range(8)]
    new = [np.zeros_like(state[0], order="F") for _ in range(8)]
    args = [_i(nx), _i(ny), _i(nz), _i(nh), _d(3.14159265358979323846), _d(0.0)]
    args += [_ptr(s) for s in state]
    for o, n in zip(old, new):
        args += [_ptr(o), _ptr(n)]
    (ref().derive3com_ if mu is None else ref().derive3com_mueta_)(*args)
    return new


def ref_ab3com(state, f_old, f_new, nx, dt):
    """Calls the reference ab3com (src/ab3com.f95:6-66; dt from COMMON)."""
    state, f_old, f_new = ([np.array(s, dtype=np.complex128, order="F") for s in x] for x in (state, f_old, f_new))
    nh, ny, nz = state[0].shape
    ref_comm().dt = dt
    new = [np.zeros_like(state[0], order="F") for _ in range(8)]
    args = [_i(nx), _i(ny), _i(nz), _i(nh), _d(3.14159265358979323846), _d(0.0)]
    args += [_ptr(s) for s in state] + [_ptr(s) for s in new]
    for o, n in zip(f_old, f_new):
        args += [_ptr(o), _ptr(n)]
    ref().ab3com_(*args)
    return new


def ref_mhd_ab2_steps(state, nx, lx, ly, lz, mu, eta, dt, nsteps, cs=10.0, threads=1, hist=None):
    """The reference's 3D AB2 loop (src/main3com.f95:536-617) over
    derive3com_mueta + ab3com, history starting at zero unless given.
    Returns (state, hist) after nsteps steps."""
    nh, ny, nz = state[0].shape
    st = np.stack([np.asarray(s, dtype=np.complex128) for s in state], axis=-1)
    st = np.asfortranarray(st)
    hs = np.zeros_like(st, order="F") if hist is None else np.asfortranarray(
        np.stack([np.asarray(s, dtype=np.complex128) for s in hist], axis=-1))
    _set_comm3(nx, ny, nz, lx, ly, lz, cs, threads, mu, eta, dt)
    ref().ref_mhd_ab2_steps(nx, ny, nz, nsteps, 1, _ptr(st), _ptr(hs))
    return ([np.asfortranarray(st[..., f]) for f in range(8)], [np.asfortranarray(hs[..., f]) for f in range(8)])
