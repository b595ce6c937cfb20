"""Grids that are not powers of two (the reference takes any Nx, Ny from
input.ini and FFTW any length): the host-array transforms and the drop-in
derive2csh / derive2css / rk42css on such grids, through the C ABI, against
numpy's FFT, the oracle and the reference routines built with FFTW 3.3.9."""
import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rfftn_ref_layout(real):
    """numpy's r2c in the reference layout (half spectrum over the first,
    fastest index)."""
    return np.asfortranarray(np.fft.rfftn(real.T).T)


@pytest.mark.parametrize("shape", [(96, 80), (100, 6), (127, 45), (6, 10), (64, 96), (2, 3), (4096, 24)])
def test_r2c_2d_any_size(shape):
    print(f"r2c {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    real = np.asfortranarray(rng.standard_normal(shape))
    got = T.fft_r2c(real)
    assert rel_l2(got, _rfftn_ref_layout(real)) < 1e-14


@pytest.mark.parametrize("shape", [(96, 80), (101, 12), (6, 10), (64, 96), (210, 7)])
def test_c2r_2d_any_size_fftw_semantics(shape):
    """A half spectrum that is not Hermitian in its kx = 0 (and even-Nx
    Nyquist) column: FFTW's c2r reads only the real parts there, as the
    oracle's c2r."""
    print(f"c2r {shape}", flush=True)
    nx, ny = shape
    rng = np.random.default_rng(nx * 7 + ny)
    spec = np.asfortranarray(rng.standard_normal((nx // 2 + 1, ny)) + 1j * rng.standard_normal((nx // 2 + 1, ny)))
    got = T.fft_c2r(spec, nx)
    want = O.c2r(spec, nx)
    assert rel_l2(got, want) < 1e-14
    assert np.array_equal(spec, np.asfortranarray(spec))


@pytest.mark.parametrize("shape", [(30, 18, 12), (16, 24, 10), (9, 8, 8), (48, 48, 40)])
def test_r2c_c2r_3d_any_size(shape):
    print(f"3d {shape}", flush=True)
    rng = np.random.default_rng(shape[0])
    real = np.asfortranarray(rng.standard_normal(shape))
    spec = T.fft_r2c(real)
    assert rel_l2(spec, _rfftn_ref_layout(real)) < 1e-14
    back = T.fft_c2r(spec, shape[0]) / np.prod(shape)
    assert rel_l2(back, real) < 1e-14


def _field(nx, ny, seed):
    """A full-spectrum random vorticity (white noise of unit variance)."""
    rng = np.random.default_rng(seed)
    return O.r2c(np.asfortranarray(rng.standard_normal((nx, ny))))


@pytest.mark.parametrize("nx,ny", [(96, 96), (100, 60), (48, 64), (64, 40), (30, 34)])
def test_derive2csh_any_grid_vs_reference(nx, ny):
    """The drop-in on grids the fused passes do not take, against the
    reference derive2csh (with FFTW) and the oracle."""
    print(f"derive2csh {nx}x{ny}", flush=True)
    w = _field(nx, ny, nx + ny)
    got = T.derive2csh(nx, ny, nx // 2 + 1, 1e-3, w)
    assert rel_l2(got, O.derive2csh(w, nx, 1e-3)) < 1e-12
    if O.ref_available():
        assert rel_l2(got, O.ref_derive2csh(w, nx, 1e-3)) < 1e-12


@pytest.mark.parametrize("nx,ny", [(96, 80), (40, 40)])
def test_derive2css_and_rk42css_any_grid_vs_oracle(nx, ny):
    print(f"derive2css/rk42css {nx}x{ny}", flush=True)
    w = _field(nx, ny, 3 * nx + ny)
    got = T.derive2css(nx, ny, nx // 2 + 1, 0.7, 1e-3, w, lam=10.0)
    assert rel_l2(got, O.derive2css(w, nx, 0.7, 1e-3, 10.0)) < 1e-12
    got = T.rk42css(nx, ny, nx // 2 + 1, 0.7, 1e-3, 1e-3, w, lam=10.0)
    assert rel_l2(got, O.rk4css_steps(w, nx, 0.7, 1e-3, 10.0, 1e-3, 1)) < 1e-13


def test_ab2_loop_any_grid_vs_oracle():
    """The drop-in derive2csh + ab2csh loop of main2csh on a 96 x 72 grid,
    20 steps from the shear layer, against the oracle's loop."""
    nx, ny, nu, dt = 96, 72, 1e-2, 1e-3
    _, w = O.ic2csh(nx, ny, d_in=0.0234375)
    ref, hist = O.ab2csh_steps(w, np.zeros_like(w, order="F"), nx, nu, dt, 20)
    old = np.zeros_like(w, order="F")
    for _ in range(20):
        f = T.derive2csh(nx, ny, nx // 2 + 1, nu, w)
        w = T.ab2csh(nx, ny, nx // 2 + 1, dt, w, old, f)
        old = f
    assert rel_l2(w, ref) < 1e-12


def _mhd_state(shape, seed):
    """rho = 1 + small, rho u and B smooth random fields, E = 1: half spectra
    in the reference layout."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x, y, z = (np.arange(n) * 2 * np.pi / n for n in shape)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")

    def field(a):
        f = np.zeros(shape)
        for _ in range(5):
            k = rng.integers(-2, 3, size=3)
            f += rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        return a * f

    reals = [1.0 + field(0.02), field(0.1), field(0.1), field(0.1), np.ones(shape), field(0.1), field(0.1),
             field(0.1)]
    return [T.fft_r2c(np.asfortranarray(r)) for r in reals]


@pytest.mark.parametrize("shape", [(24, 20, 18), (12, 16, 10), (32, 32, 24), (48, 40, 36)])
def test_derive3com_any_grid_vs_oracle(shape):
    """The drop-in derive3com on grids the fused passes do not take, with
    mu != eta and a box that is not 2 pi, against the oracle."""
    print(f"derive3com {shape}", flush=True)
    nx, ny, nz = shape
    st = _mhd_state(shape, nx + ny + nz)
    lx, ly, lz = 2 * np.pi, 3 * np.pi, 2.5 * np.pi
    got = T.derive3com(nx, ny, nz, nx // 2 + 1, T.MhdParams(lx, ly, lz, 1e-2, 3e-2, 10.0, 0.0), st)
    ref = O.derive3com(st, nx, lx, ly, lz, 1e-2, 3e-2, 10.0)
    for f in (0, 1, 2, 3, 5, 6, 7):
        assert rel_l2(got[f], ref[f]) < 1e-12, f
    assert np.all(got[4] == 0)


COH = dict(pi=np.pi, lx=2 * np.pi, ly=3 * np.pi, mux=2e-3, muy=3e-3, cs=3.77)
COM = dict(pi=np.pi, lx=2 * np.pi, ly=3 * np.pi, mu=1e-3, eta=2e-3, cs=2.0)


@pytest.mark.parametrize("nx,ny", [(96, 60), (24, 30), (64, 48), (40, 512)])
def test_derive2coh_any_grid_vs_oracle(nx, ny):
    print(f"derive2coh {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx * ny)
    reals = (1.0 + 0.02 * rng.standard_normal((nx, ny)), 0.1 * rng.standard_normal((nx, ny)),
             0.1 * rng.standard_normal((nx, ny)))
    st = [O.r2c(np.asfortranarray(a)) for a in reals]
    prm = T.CohParams(COH["pi"], COH["lx"], COH["ly"], COH["mux"], COH["muy"], COH["cs"], 0.0)
    got = T.derive2coh(nx, ny, nx // 2 + 1, prm, st)
    ref = O.derive2coh(st, nx, **COH)
    for k in range(3):
        assert rel_l2(got[k], ref[k]) < 1e-12, k


@pytest.mark.parametrize("nx,ny", [(96, 60), (24, 30), (64, 48)])
def test_derive2com_any_grid_with_forcing_vs_oracle(nx, ny):
    print(f"derive2com {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx + 7 * ny)
    reals = [1.0 + 0.02 * rng.standard_normal((nx, ny))] + [0.1 * rng.standard_normal((nx, ny)) for _ in range(2)]
    reals += [2.0 + 0.05 * rng.standard_normal((nx, ny))] + [0.1 * rng.standard_normal((nx, ny)) for _ in range(2)]
    st = [O.r2c(np.asfortranarray(a)) for a in reals]
    prm = T.ComParams(COM["pi"], COM["lx"], COM["ly"], COM["mu"], COM["eta"], COM["cs"], 1.0, 0.3, 0.0)
    got = T.derive2com(nx, ny, nx // 2 + 1, prm, st)
    ref = O.derive2com(st, nx, kf=1.0, a=0.3, **COM)
    for k in range(6):
        assert rel_l2(got[k], ref[k]) < 1e-12, k
