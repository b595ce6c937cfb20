"""tara_amd -- MI355X-native (gfx950) pseudo-spectral hot path of TARA.

Python mirror of the reference's hot-path interface, over the C ABI of
``libtara_hip.so`` (include/tara_hip.h).  Function names, argument meaning and
array layout follow the Fortran routines they replace:

  derive2csh(Nx, Ny, Nh, nu, omegak)            src/derive2csh.f95:3
  ab2csh(Nx, Ny, Nh, dt, omegak, f_old, f_new)  src/ab2csh.f95:3
  derive2css(Nx, Ny, Nh, alpha, nu, omegak, lam)   src/derive2css.f95:6
  ab2css / rk42css                              src/ab2css.f95:6, src/rk42css.f95:6
  derive2coh / ab2coh / euler2coh / rk4coh_field / pccoh_field
                                                src/derive2coh.f95:5, ab2coh.f95:5, euler2coh.f95:5,
                                                rk4nk2coh.f95:6, pcnk2coh.f95:6
  derive2com / ab2com                           src/derive2com.f95:5, src/ab2com.f95:6
  derive3com(Nx, Ny, Nz, Nh, params, state)     src/derive3com.f95:5
  ab3com(Nx, Ny, Nz, Nh, dt, state, f_old, f_new)  src/ab3com.f95:6
  fft_r2c / fft_c2r                             dfftw_plan_dft_{r2c,c2r}_{2d,3d} + execute

Spectral arrays are complex128 half spectra (Nh, Ny[, Nz]) in Fortran order
(numpy ``order='F'``).  Everything computes on the GPU; there is no CPU
fallback -- a missing or unusable library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TARA_HIP_LIB") or os.path.join(_HERE, "libtara_hip.so")
_dp = C.POINTER(C.c_double)
_lib = None

SCHEME_AB2 = 1
SCHEME_RK4 = 2


ERR_NONFINITE = -6


class TaraError(RuntimeError):
    """A nonzero status of the C ABI; .code holds it (e.g. ERR_NONFINITE)."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


class MhdParams(C.Structure):
    """tara_mhd_params: physical box (input.ini value times pi), diffusion
    coefficients, isothermal sound speed and time step."""
    _fields_ = [(n, C.c_double) for n in ("lx", "ly", "lz", "mu", "eta", "cs", "dt")]


class CohParams(C.Structure):
    """tara_coh_params: the pi argument, COMMON Lx, Ly (times pi), mux, muy, CS, dt."""
    _fields_ = [(n, C.c_double) for n in ("pi", "lx", "ly", "mux", "muy", "cs", "dt")]


class ComParams(C.Structure):
    """tara_com_params: the pi argument, COMMON Lx, Ly, mu, eta, CS, kf, A, dt."""
    _fields_ = [(n, C.c_double) for n in ("pi", "lx", "ly", "mu", "eta", "cs", "kf", "a", "dt")]


def lib():
    """Loads libtara_hip.so (built by ``python -m tara_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TaraError(f"{LIB_PATH} not built: run `python -m tara_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.tara_last_error.restype = C.c_char_p
        for n in ("tara_fft_r2c_1d", "tara_fft_c2r_1d"):
            getattr(L, n).argtypes = [C.c_int, _dp, _dp]
        for n in ("tara_fft_r2c_2d", "tara_fft_c2r_2d"):
            getattr(L, n).argtypes = [C.c_int, C.c_int, _dp, _dp]
        for n in ("tara_fft_r2c_3d", "tara_fft_c2r_3d"):
            getattr(L, n).argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.tara_derive2csh.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.tara_ab2csh.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.tara_csh_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_void_p)]
        L.tara_derive2css.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.tara_ab2css.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.tara_rk42css.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 4 + [_dp, _dp]
        L.tara_css_create.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4 + [C.c_int, C.POINTER(C.c_void_p)]
        L.tara_csh_upload.argtypes = [C.c_void_p, _dp]
        L.tara_csh_download.argtypes = [C.c_void_p, _dp]
        L.tara_csh_advance.argtypes = [C.c_void_p, C.c_int]
        L.tara_csh_sync.argtypes = [C.c_void_p]
        L.tara_csh_destroy.argtypes = [C.c_void_p]
        L.tara_csh_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.tara_csh_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long), _dp]
        L.tara_csh_num_xshapes.argtypes = []
        L.tara_csh_fields.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.tara_csh_record.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.tara_csh_ic_shear.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.tara_csh_ic_random.argtypes = [C.c_void_p, C.c_uint, C.c_double, C.c_double]
        L.tara_philox4x32_10.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
        L.tara_csh_tune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), _dp]
        L.tara_csh_set_xshape.argtypes = [C.c_void_p, C.c_int]
        L.tara_gemm.argtypes = ([C.c_int, C.c_int, C.c_long, C.c_long, C.c_long, C.c_void_p, C.c_long, C.c_long,
                                 C.c_void_p, C.c_long, C.c_long, C.c_void_p, C.c_long, C.c_long,
                                 C.c_int, C.c_long, C.c_long, C.c_long, C.c_int, C.c_long, C.c_long, C.c_long,
                                 C.c_float])
        L.tara_ens_create.argtypes = [C.c_int, C.POINTER(C.c_int), _dp, C.c_double, C.c_int, C.POINTER(C.c_void_p)]
        L.tara_ens_upload.argtypes = [C.c_void_p, C.c_int, _dp]
        L.tara_ens_download.argtypes = [C.c_void_p, C.c_int, _dp]
        L.tara_ens_advance.argtypes = [C.c_void_p, C.c_int]
        for n in ("tara_ens_sync", "tara_ens_destroy"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.tara_ens_use_graph.argtypes = [C.c_void_p, C.c_int]
        L.tara_ens_tune.argtypes = [C.c_void_p, C.c_int]
        L.tara_ens_ic_random.argtypes = [C.c_void_p, C.c_int, C.c_uint, C.c_double, C.c_double]
        L.tara_ens_graph_nodes.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.tara_ens_nonfinite.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.tara_ens_set_omega.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tara_ens_get_omega.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _fp = C.POINTER(C.c_float)
        L.tara_fno_create.argtypes = [C.c_int] * 7 + [C.c_uint, C.POINTER(C.c_void_p)]
        L.tara_fno_num_params.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
        L.tara_fno_get_params.argtypes = [C.c_void_p, _fp]
        L.tara_fno_set_params.argtypes = [C.c_void_p, _fp]
        L.tara_fno_forward.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.tara_fno_grads.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, _fp, _fp]
        L.tara_fno_train_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float, _fp]
        L.tara_fno_rel_err.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, _fp]
        L.tara_fno_gather_pairs.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long), C.c_long, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p]
        L.tara_fno_destroy.argtypes = [C.c_void_p]
        L.tara_fno_rollout.argtypes = ([C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_void_p), C.c_void_p, C.c_int, C.c_float, C.c_int, _fp,
                                        C.POINTER(C.c_int)])
        P3, P6 = C.c_void_p * 3, C.c_void_p * 6
        L.tara_derive2coh.argtypes = [C.c_int] * 3 + [C.POINTER(CohParams), P3, P3]
        L.tara_ab2coh.argtypes = [C.c_int] * 3 + [C.c_double, P3, P3, P3, P3]
        L.tara_euler2coh.argtypes = [C.c_int] * 3 + [C.c_double, P3, P3, P3]
        for n in ("tara_rk4coh_field", "tara_pccoh_field"):
            getattr(L, n).argtypes = [C.c_int] * 3 + [C.c_double, _dp, _dp, _dp]
        L.tara_derive2com.argtypes = [C.c_int] * 3 + [C.POINTER(ComParams), P6, P6]
        L.tara_ab2com.argtypes = [C.c_int] * 3 + [C.c_double, P6, P6, P6, P6]
        L.tara_coh_create.argtypes = [C.c_int, C.c_int, C.POINTER(CohParams), C.POINTER(C.c_void_p)]
        L.tara_com_create.argtypes = [C.c_int, C.c_int, C.POINTER(ComParams), C.POINTER(C.c_void_p)]
        L.tara_f2_upload.argtypes = [C.c_void_p, C.c_void_p]
        L.tara_f2_download.argtypes = [C.c_void_p, C.c_void_p]
        L.tara_f2_advance.argtypes = [C.c_void_p, C.c_int]
        for n in ("tara_f2_sync", "tara_f2_destroy"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.tara_f2_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.tara_f2_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long), _dp]
        P8 = C.c_void_p * 8
        L.tara_derive3com.argtypes = [C.c_int] * 4 + [C.POINTER(MhdParams), P8, P8]
        L.tara_ab3com.argtypes = [C.c_int] * 4 + [C.c_double, P8, P8, P8, P8]
        L.tara_mhd_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(MhdParams), C.POINTER(C.c_void_p)]
        L.tara_mhd_upload.argtypes = [C.c_void_p, P8]
        L.tara_mhd_download.argtypes = [C.c_void_p, P8]
        L.tara_mhd_advance.argtypes = [C.c_void_p, C.c_int]
        for n in ("tara_mhd_sync", "tara_mhd_destroy"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.tara_mhd_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.tara_mhd_ic_abc.argtypes = [C.c_void_p] + [C.c_double] * 8
        L.tara_mhd_ic_random.argtypes = [C.c_void_p, C.c_uint] + [C.c_double] * 3
        L.tara_mhd_energy.argtypes = [C.c_void_p, _dp]
        L.tara_csh_advance_tracers.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                               C.c_int, _dp, _dp]
        L.tara_mhd_record_energy.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.tara_derive2csh_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
        L.tara_mhd_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long), _dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise TaraError(f"tara error {rc}: {lib().tara_last_error().decode()}", rc)


def _p(a):
    return a.ctypes.data_as(_dp)


def _spec(a):
    return np.asfortranarray(a, dtype=np.complex128)


def _real(a):
    return np.asfortranarray(a, dtype=np.float64)


# ------------------------------------------------------------------ transforms
def fft_r2c(real):
    """r2c of a (Nx, Ny[, Nz]) real field -> (Nh, Ny[, Nz]) half spectrum (FFTW semantics)."""
    real = _real(real)
    if real.ndim == 2:
        nx, ny = real.shape
        out = np.zeros((nx // 2 + 1, ny), np.complex128, order="F")
        _check(lib().tara_fft_r2c_2d(nx, ny, _p(real), _p(out)))
    else:
        nx, ny, nz = real.shape
        out = np.zeros((nx // 2 + 1, ny, nz), np.complex128, order="F")
        _check(lib().tara_fft_r2c_3d(nx, ny, nz, _p(real), _p(out)))
    return out


def fft_c2r(spec, nx):
    """Unnormalised c2r (FFTW semantics); the input is left intact."""
    spec = _spec(spec)
    if spec.ndim == 2:
        ny = spec.shape[1]
        out = np.zeros((nx, ny), order="F")
        _check(lib().tara_fft_c2r_2d(nx, ny, _p(spec), _p(out)))
    else:
        ny, nz = spec.shape[1:]
        out = np.zeros((nx, ny, nz), order="F")
        _check(lib().tara_fft_c2r_3d(nx, ny, nz, _p(spec), _p(out)))
    return out


# ------------------------------------------------- 2D incompressible (csh)
def derive2csh(Nx, Ny, Nh, nu, omegak):
    """dt_omegak_new = derive2csh(...) (src/derive2csh.f95:3)."""
    w = _spec(omegak)
    assert w.shape == (Nh, Ny)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_derive2csh(Nx, Ny, Nh, nu, _p(w), _p(out)))
    return out


def derive2csh_noalloc(Nx, Ny, Nh, nu, omegak_dev, dt_omegak_new_dev, stream=0):
    """derive2csh on device arrays (pointers to (Nh, Ny) complex128 half
    spectra in the reference layout) without any device allocation, enqueued
    on `stream` (a hipStream_t as an int; 0 = the null stream).  Grids up to
    1024 x 1024; calls that overlap on one GPU race (they share its scratch)."""
    _check(lib().tara_derive2csh_noalloc(Nx, Ny, Nh, nu, omegak_dev, dt_omegak_new_dev, stream or None))


def ab2csh(Nx, Ny, Nh, dt, omegak, dt_omegak_old, dt_omegak_new):
    """omegak_new = ab2csh(...) (src/ab2csh.f95:3)."""
    w, fo, fn = _spec(omegak), _spec(dt_omegak_old), _spec(dt_omegak_new)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_ab2csh(Nx, Ny, Nh, dt, _p(w), _p(out), _p(fo), _p(fn)))
    return out


# lambda of the screened form when the caller gives none: main2css.f95:87
CSS_LAMBDA = 10.0


def derive2css(Nx, Ny, Nh, alpha, nu, omegak, lam=CSS_LAMBDA):
    """dt_omegak_new = derive2css(...) (src/derive2css.f95:6); Lx = Ly = 2 pi."""
    w = _spec(omegak)
    assert w.shape == (Nh, Ny)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_derive2css(Nx, Ny, Nh, alpha, nu, lam, _p(w), _p(out)))
    return out


def ab2css(Nx, Ny, Nh, dt, omegak, dt_omegak_old, dt_omegak_new):
    """omegak_new = ab2css(...) (src/ab2css.f95:6)."""
    w, fo, fn = _spec(omegak), _spec(dt_omegak_old), _spec(dt_omegak_new)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_ab2css(Nx, Ny, Nh, dt, _p(w), _p(out), _p(fo), _p(fn)))
    return out


def rk42css(Nx, Ny, Nh, alpha, nu, dt, omegak, lam=CSS_LAMBDA):
    """omegak_new = one classic RK4 step of derive2css (the intent of src/rk42css.f95:6)."""
    w = _spec(omegak)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_rk42css(Nx, Ny, Nh, alpha, nu, lam, dt, _p(w), _p(out)))
    return out


class CshSolver:
    """Device-resident 2D vorticity solver (the main2csh.f95:114-126 loop on the
    GPU).  With ``screened=True`` it steps derive2css's equations instead
    (alpha, lam as in derive2css)."""

    def __init__(self, nx, ny, nu, dt, scheme=SCHEME_AB2, screened=False, alpha=1.0, lam=CSS_LAMBDA):
        self.nx, self.ny, self.nh, self.dt = nx, ny, nx // 2 + 1, dt
        h = C.c_void_p()
        if screened:
            _check(lib().tara_css_create(nx, ny, alpha, nu, lam, dt, scheme, C.byref(h)))
        else:
            _check(lib().tara_csh_create(nx, ny, nu, dt, scheme, C.byref(h)))
        self._h = h

    def upload(self, omegak):
        w = _spec(omegak)
        assert w.shape == (self.nh, self.ny)
        _check(lib().tara_csh_upload(self._h, _p(w)))

    def download(self):
        out = np.zeros((self.nh, self.ny), np.complex128, order="F")
        _check(lib().tara_csh_download(self._h, _p(out)))
        return out

    def advance(self, nsteps):
        _check(lib().tara_csh_advance(self._h, nsteps))

    def sync(self):
        _check(lib().tara_csh_sync(self._h))

    def set_timing(self, on):
        _check(lib().tara_csh_set_timing(self._h, 1 if on else 0))

    def kernel_stats(self, nclass=3):
        counts = (C.c_long * nclass)()
        ms = (C.c_double * nclass)()
        _check(lib().tara_csh_kernel_stats(self._h, nclass, counts, ms))
        return list(counts), list(ms)

    def tune(self, reps=5):
        """Times every x-pass shape on this grid, keeps the fastest; returns
        (chosen shape, per-shape ms with -1 for shapes that do not run)."""
        n = lib().tara_csh_num_xshapes()
        ms = (C.c_double * n)()
        chosen = C.c_int()
        _check(lib().tara_csh_tune(self._h, reps, C.byref(chosen), ms))
        return chosen.value, list(ms)

    def set_xshape(self, shape):
        _check(lib().tara_csh_set_xshape(self._h, shape))

    def fields(self):
        """Real-space (omega, ux, uy), each (Nx, Ny), as src/main2csh.f95:128-178 forms them."""
        out = [np.zeros((self.nx, self.ny), order="F") for _ in range(3)]
        _check(lib().tara_csh_fields(self._h, *[_p(a) for a in out]))
        return out

    def record(self, nframes, every):
        """Advances nframes * every steps; returns omegak after every `every` steps."""
        frames = np.zeros((nframes, self.ny, self.nh), np.complex128)
        _check(lib().tara_csh_record(self._h, nframes, every, frames.ctypes.data_as(_dp)))
        return [np.asfortranarray(f.T) for f in frames]

    def advance_tracers(self, nsteps, xp, yp, lx=2 * np.pi, ly=2 * np.pi, dt=None, periodic=False):
        """Steps nsteps times, moving the tracer particles (xp, yp: float64
        arrays in [0, lx] x [0, ly], updated in place) after each step as
        src/main2csh.f95:187-265; periodic=False reads the interpolation
        corners as the driver addresses its arrays, True takes their periodic
        images."""
        assert xp.dtype == np.float64 and yp.dtype == np.float64 and xp.size == yp.size
        assert xp.flags.c_contiguous and yp.flags.c_contiguous
        _check(lib().tara_csh_advance_tracers(self._h, nsteps, self.dt if dt is None else dt, lx, ly,
                                              int(bool(periodic)), xp.size, _p(xp), _p(yp)))

    def ic_shear(self, lx_in=2.0, ly_in=2.0, w0=25.0, m0=2.0, d_in=3.0 / 128.0):
        """The reference's shear layer generated on the device (history zeroed)."""
        _check(lib().tara_csh_ic_shear(self._h, lx_in, ly_in, w0, m0, d_in))

    def ic_random(self, seed, k0=4.0, energy=1.0):
        """Philox random-phase vorticity generated on the device (history zeroed)."""
        _check(lib().tara_csh_ic_random(self._h, seed, k0, energy))

    def close(self):
        if self._h:
            lib().tara_csh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GEMM_BF16, GEMM_F32 = 0, 1
GEMM_STORE_F32, GEMM_ACCUM_F32, GEMM_STORE_BF16 = 0, 1, 2


def gemm(dtype, epi, m, n, k, a, sam, sak, b, sbk, sbn, c, scm, scn,
         nb1=1, sa1=0, sb1=0, sc1=0, nb2=1, sa2=0, sb2=0, sc2=0, alpha=1.0):
    """tara_gemm on device pointers (ints), strides in elements."""
    _check(lib().tara_gemm(dtype, epi, m, n, k, a, sam, sak, b, sbk, sbn, c, scm, scn,
                           nb1, sa1, sb1, sc1, nb2, sa2, sb2, sc2, alpha))


def _dptr(x):
    """Device address of a torch device tensor, or an int taken as one."""
    return x.data_ptr() if hasattr(x, "data_ptr") else int(x)


FNO_BF16, FNO_FP32 = 0, 1


class FNO:
    """Fourier-neural-operator surrogate on the matrix cores (tara_fno_*):
    omega(t) -> omega(t + T) for batches of n x n real-space vorticity fields in
    device float32 memory (torch tensors or device addresses)."""

    def __init__(self, n, width=64, mx=16, my=32, layers=4, max_batch=16, precision=FNO_BF16, seed=0):
        self.n, self.width, self.mx, self.my, self.layers = n, width, mx, my, layers
        self.max_batch = max_batch
        h = C.c_void_p()
        _check(lib().tara_fno_create(n, width, mx, my, layers, max_batch, precision, seed, C.byref(h)))
        self._h = h
        k = C.c_long()
        _check(lib().tara_fno_num_params(h, C.byref(k)))
        self.nparams = k.value

    def get_params(self):
        out = np.zeros(self.nparams, np.float32)
        _check(lib().tara_fno_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        _check(lib().tara_fno_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, batch, omega, out):
        _check(lib().tara_fno_forward(self._h, batch, _dptr(omega), _dptr(out)))

    def grads(self, batch, omega, target):
        """(loss, gradient as float32 numpy) of the batch-mean relative L2 error."""
        loss = C.c_float()
        g = np.zeros(self.nparams, np.float32)
        _check(lib().tara_fno_grads(self._h, batch, _dptr(omega), _dptr(target), C.byref(loss),
                                    g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def train_step(self, batch, omega, target, lr):
        loss = C.c_float()
        _check(lib().tara_fno_train_step(self._h, batch, _dptr(omega), _dptr(target), lr, C.byref(loss)))
        return loss.value

    def gather_pairs(self, idx, inp, tgt, xb, yb):
        """xb[b] = inp[idx[b]], yb[b] = tgt[idx[b]] on the device (inp, tgt:
        (npairs, n, n) float32 device tensors; idx: host integer array)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        npairs = inp.shape[0]
        _check(lib().tara_fno_gather_pairs(self._h, idx.size, idx.ctypes.data_as(C.POINTER(C.c_long)), npairs,
                                           _dptr(inp), _dptr(tgt), _dptr(xb), _dptr(yb)))

    def rel_err(self, batch, a, ref):
        out = np.zeros(batch, np.float32)
        _check(lib().tara_fno_rel_err(self._h, batch, _dptr(a), _dptr(ref),
                                      out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def rollout(self, batch, traj, nsteps, every=10, checkers=(), solver_steps=100, tol=0.05, graph=True,
                anchor=None):
        """Surrogate rollout of traj[0] (device float32 [nsteps+1][batch][n][n])
        with FP64 checks every `every` steps by the Ensemble objects in
        `checkers`: nsteps // every of them with `batch` runs each (checks start
        as their inputs appear), or one with (nsteps // every) * batch runs (all
        checks advanced together after the rollout).  `anchor` (an Ensemble of
        `batch` runs of the same solver) carries the FP64 solution that takes
        the trajectory over at a failed check.  Returns (check errors
        [nchk][batch], number of solver take-overs)."""
        nens = len(checkers)
        nchk = nsteps // every
        hs = (C.c_void_p * max(nens, 1))(*[c._h for c in checkers])
        err = np.zeros(max(nchk * batch, 1), np.float32)
        nf = C.c_int()
        _check(lib().tara_fno_rollout(self._h, batch, _dptr(traj), nsteps, every, nens, hs,
                                      anchor._h if anchor is not None else None, solver_steps, tol,
                                      1 if graph else 0, err.ctypes.data_as(C.POINTER(C.c_float)), C.byref(nf)))
        return err[:nchk * batch].reshape(nchk, batch), nf.value

    def close(self):
        if self._h:
            lib().tara_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def philox4x32_10(ctr, key):
    """Host evaluation of the generator behind CshSolver.ic_random."""
    c = (C.c_uint * 4)(*ctr)
    k = (C.c_uint * 2)(*key)
    o = (C.c_uint * 4)()
    _check(lib().tara_philox4x32_10(c, k, o))
    return list(o)


class Ensemble:
    """Batched ensemble of independent 2D incompressible runs (derive2csh
    physics): run r has an n[r] x n[r] grid and viscosity nu[r]; all share dt and
    the scheme.  Buckets of one grid size advance in batched launches; a step is
    replayed from a captured hipGraph."""

    def __init__(self, n, nu, dt, scheme=SCHEME_RK4):
        self.n = [int(v) for v in n]
        arr_n = (C.c_int * len(self.n))(*self.n)
        arr_nu = (C.c_double * len(self.n))(*[float(v) for v in nu])
        h = C.c_void_p()
        _check(lib().tara_ens_create(len(self.n), arr_n, arr_nu, dt, scheme, C.byref(h)))
        self._h = h

    def upload(self, run, omegak):
        n = self.n[run]
        w = _spec(omegak)
        assert w.shape == (n // 2 + 1, n)
        _check(lib().tara_ens_upload(self._h, run, _p(w)))

    def download(self, run):
        n = self.n[run]
        out = np.zeros((n // 2 + 1, n), np.complex128, order="F")
        _check(lib().tara_ens_download(self._h, run, _p(out)))
        return out

    def advance(self, nsteps):
        _check(lib().tara_ens_advance(self._h, nsteps))

    def sync(self):
        _check(lib().tara_ens_sync(self._h))

    def use_graph(self, on):
        _check(lib().tara_ens_use_graph(self._h, 1 if on else 0))

    def ic_random(self, run, seed, k0=4.0, energy=1.0):
        """Random-phase vorticity of CshSolver.ic_random for one run."""
        _check(lib().tara_ens_ic_random(self._h, run, seed, k0, energy))

    def set_omega(self, run, omega):
        """Real-space vorticity (device float32, n x n, x fastest) into one run."""
        _check(lib().tara_ens_set_omega(self._h, run, _dptr(omega)))

    def get_omega(self, run, out):
        """One run's real-space vorticity into device float32 memory (asynchronous
        until sync())."""
        _check(lib().tara_ens_get_omega(self._h, run, _dptr(out)))

    def nonfinite(self):
        """Per-run flags (list of bool) of runs whose state became NaN or Inf."""
        flags = (C.c_int * len(self.n))()
        rc = lib().tara_ens_nonfinite(self._h, flags)
        if rc < 0:
            _check(rc)
        return [bool(f) for f in flags]

    def graph_nodes(self):
        """(nodes, kernel nodes) of the captured step graph; (0, 0) before capture."""
        n, k = C.c_long(), C.c_long()
        _check(lib().tara_ens_graph_nodes(self._h, C.byref(n), C.byref(k)))
        return n.value, k.value

    def tune(self, reps=3):
        _check(lib().tara_ens_tune(self._h, reps))

    def close(self):
        if self._h:
            lib().tara_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------- 2D compressible families (coh, com)
def _pn(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _specs(arrs, shape):
    out = [_spec(a) for a in arrs]
    assert all(a.shape == shape for a in out)
    return out


def derive2coh(Nx, Ny, Nh, params, state):
    """[dn_dt_new, dux_dt_new, duy_dt_new] = derive2coh(...) (src/derive2coh.f95:5)."""
    st = _specs(state, (Nh, Ny))
    out = [np.zeros((Nh, Ny), np.complex128, order="F") for _ in range(3)]
    _check(lib().tara_derive2coh(Nx, Ny, Nh, C.byref(params), _pn(st), _pn(out)))
    return out


def ab2coh(Nx, Ny, Nh, dt, state, f_old, f_new):
    st, fo, fn = (_specs(x, (Nh, Ny)) for x in (state, f_old, f_new))
    out = [np.zeros((Nh, Ny), np.complex128, order="F") for _ in range(3)]
    _check(lib().tara_ab2coh(Nx, Ny, Nh, dt, _pn(st), _pn(out), _pn(fo), _pn(fn)))
    return out


def euler2coh(Nx, Ny, Nh, dt, state, slope):
    st, sl = (_specs(x, (Nh, Ny)) for x in (state, slope))
    out = [np.zeros((Nh, Ny), np.complex128, order="F") for _ in range(3)]
    _check(lib().tara_euler2coh(Nx, Ny, Nh, dt, _pn(st), _pn(out), _pn(sl)))
    return out


def rk4coh_field(Nx, Ny, Nh, dt, field, slope):
    """The field update of rk4_nk2coh / rk4_ukx2coh / rk4_uky2coh."""
    w, s_ = _spec(field), _spec(slope)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_rk4coh_field(Nx, Ny, Nh, dt, _p(w), _p(out), _p(s_)))
    return out


def pccoh_field(Nx, Ny, Nh, dt, field, slope):
    """The field update of pc_nk2coh / pc_ukx2coh / pc_uky2coh."""
    w, s_ = _spec(field), _spec(slope)
    out = np.zeros_like(w, order="F")
    _check(lib().tara_pccoh_field(Nx, Ny, Nh, dt, _p(w), _p(out), _p(s_)))
    return out


def derive2com(Nx, Ny, Nh, params, state):
    """The six d_*_dt_new of derive2com(...) (src/derive2com.f95:5)."""
    st = _specs(state, (Nh, Ny))
    out = [np.zeros((Nh, Ny), np.complex128, order="F") for _ in range(6)]
    _check(lib().tara_derive2com(Nx, Ny, Nh, C.byref(params), _pn(st), _pn(out)))
    return out


def ab2com(Nx, Ny, Nh, dt, state, f_old, f_new):
    st, fo, fn = (_specs(x, (Nh, Ny)) for x in (state, f_old, f_new))
    out = [np.zeros((Nh, Ny), np.complex128, order="F") for _ in range(6)]
    _check(lib().tara_ab2com(Nx, Ny, Nh, dt, _pn(st), _pn(out), _pn(fo), _pn(fn)))
    return out


class F2Solver:
    """Device-resident AB2 stepping of derive2coh (params: CohParams) or
    derive2com (params: ComParams)."""

    def __init__(self, nx, ny, params):
        self.nx, self.ny, self.nh = nx, ny, nx // 2 + 1
        h = C.c_void_p()
        if isinstance(params, CohParams):
            self.nf = 3
            _check(lib().tara_coh_create(nx, ny, C.byref(params), C.byref(h)))
        else:
            self.nf = 6
            _check(lib().tara_com_create(nx, ny, C.byref(params), C.byref(h)))
        self._h = h

    def upload(self, state):
        st = _specs(state, (self.nh, self.ny))
        assert len(st) == self.nf
        _check(lib().tara_f2_upload(self._h, _pn(st)))

    def download(self):
        out = [np.zeros((self.nh, self.ny), np.complex128, order="F") for _ in range(self.nf)]
        _check(lib().tara_f2_download(self._h, _pn(out)))
        return out

    def advance(self, nsteps):
        _check(lib().tara_f2_advance(self._h, nsteps))

    def sync(self):
        _check(lib().tara_f2_sync(self._h))

    def set_timing(self, on):
        _check(lib().tara_f2_set_timing(self._h, 1 if on else 0))

    def kernel_stats(self, nclass=3):
        counts = (C.c_long * nclass)()
        ms = (C.c_double * nclass)()
        _check(lib().tara_f2_kernel_stats(self._h, nclass, counts, ms))
        return list(counts), list(ms)

    def close(self):
        if self._h:
            lib().tara_f2_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------- 3D compressible MHD (3com)
def _p8(arrs):
    return (C.c_void_p * 8)(*[a.ctypes.data for a in arrs])


def derive3com(Nx, Ny, Nz, Nh, params, state):
    """8 RHS spectra d_*_dt_new = derive3com(...) (src/derive3com.f95:5);
    state = [rho, rho_ux, rho_uy, rho_uz, E, Bx, By, Bz], each (Nh, Ny, Nz)."""
    st = [_spec(a) for a in state]
    out = [np.zeros_like(st[0], order="F") for _ in range(8)]
    _check(lib().tara_derive3com(Nx, Ny, Nz, Nh, C.byref(params), _p8(st), _p8(out)))
    return out


def ab3com(Nx, Ny, Nz, Nh, dt, state, f_old, f_new):
    """state_new = ab3com(...) (src/ab3com.f95:6); E is copied unchanged."""
    st, fo, fn = ([_spec(a) for a in x] for x in (state, f_old, f_new))
    out = [np.zeros_like(st[0], order="F") for _ in range(8)]
    _check(lib().tara_ab3com(Nx, Ny, Nz, Nh, dt, _p8(st), _p8(out), _p8(fo), _p8(fn)))
    return out


class MhdSolver:
    """Device-resident AB2 stepping of derive3com + ab3com (main3com.f95:536-617)."""

    def __init__(self, nx, ny, nz, params):
        self.shape = (nx // 2 + 1, ny, nz)
        h = C.c_void_p()
        _check(lib().tara_mhd_create(nx, ny, nz, C.byref(params), C.byref(h)))
        self._h = h

    def upload(self, state):
        st = [_spec(a) for a in state]
        assert all(a.shape == self.shape for a in st)
        _check(lib().tara_mhd_upload(self._h, _p8(st)))

    def download(self):
        out = [np.zeros(self.shape, np.complex128, order="F") for _ in range(8)]
        _check(lib().tara_mhd_download(self._h, _p8(out)))
        return out

    def advance(self, nsteps):
        _check(lib().tara_mhd_advance(self._h, nsteps))

    def ic_abc(self, rho0=1.0, u0=0.1, b0=0.1, p0=1.0, a=1.0, b=1.0, c=1.0, kf=1.0):
        """The reference's ABC-flow initial state (src/main3com.f95:290-324) on the device."""
        _check(lib().tara_mhd_ic_abc(self._h, rho0, u0, b0, p0, a, b, c, kf))

    def ic_random(self, seed=1, k0=3.0, e_kin=1.0, e_mag=0.5):
        """Random-phase solenoidal u and B on the device (tara_mhd_ic_random)."""
        _check(lib().tara_mhd_ic_random(self._h, seed, k0, e_kin, e_mag))

    ENERGY_COLUMNS = ("Pressure", "Energy", "T_Energy", "B_Field", "C1", "C2", "Hf", "Hm", "HT", "div_B_Tot")

    def energy(self):
        """The ten Energy.dat averages of src/main3com.f95:1070 for the current state."""
        out = np.zeros(10)
        _check(lib().tara_mhd_energy(self._h, _p(out)))
        return out

    def record_energy(self, nrec, every=1):
        """Advances nrec x every steps; returns the averages after every `every` steps, (nrec, 10)."""
        rows = np.zeros((nrec, 10))
        _check(lib().tara_mhd_record_energy(self._h, nrec, every, _p(rows)))
        return rows

    def sync(self):
        _check(lib().tara_mhd_sync(self._h))

    def set_timing(self, on):
        _check(lib().tara_mhd_set_timing(self._h, 1 if on else 0))

    def kernel_stats(self, nclass=4):
        counts = (C.c_long * nclass)()
        ms = (C.c_double * nclass)()
        _check(lib().tara_mhd_kernel_stats(self._h, nclass, counts, ms))
        return list(counts), list(ms)

    def close(self):
        if self._h:
            lib().tara_mhd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
