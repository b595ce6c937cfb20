"""Device tracer particles (src/main2csh.f95:187-265) through the C ABI:
against the oracle's restatement step by step over a run of the driver's
shear-layer flow, and against the closed form of a steady shear flow."""
import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import rel_l2

pytestmark = pytest.mark.gpu
L = 2.0 * np.pi


def _pdist(a, b, period):
    return np.abs(np.mod(a - b + 0.5 * period, period) - 0.5 * period)


def _particles(nx, ny, lx, ly, n, seed):
    """Uniform particles plus ones in the last column, the last row, the
    corner cell and on the box's edges (x = 0, x = Lx)."""
    rng = np.random.default_rng(seed)
    xp, yp = rng.uniform(0, lx, n), rng.uniform(0, ly, n)
    dx, dy = lx / nx, ly / ny
    xp[:16] = lx - dx * rng.uniform(0.01, 0.99, 16)
    yp[16:32] = ly - dy * rng.uniform(0.01, 0.99, 16)
    xp[32:36] = lx - dx * rng.uniform(0.01, 0.99, 4)
    yp[32:36] = ly - dy * rng.uniform(0.01, 0.99, 4)
    xp[36], xp[37] = 0.0, lx
    return np.ascontiguousarray(xp), np.ascontiguousarray(yp)


@pytest.mark.parametrize("periodic", [False, True])
def test_tracers_vs_oracle_shear_layer(periodic):
    """The driver's run: shear-layer IC (src/main2csh.f95:84-106), AB2 steps,
    after each step the velocity of the new state and one particle update.
    300 steps at 128^2 against the oracle's derive2csh/ab2csh steps and its
    restatement of the driver's velocity and tracer code; positions to 1e-10
    (periodic distance), the state to 1e-12."""
    nx = ny = 128
    nu, dt, steps = 1e-3, 1e-3, 300
    print(f"tracers 128^2 x {steps} periodic={periodic}", flush=True)
    _, w = O.ic2csh(nx, ny, d_in=3.0 / 128.0)
    xp, yp = _particles(nx, ny, L, L, 400, seed=7)
    xd, yd = xp.copy(), yp.copy()
    s = T.CshSolver(nx, ny, nu, dt, T.SCHEME_AB2)
    s.upload(w)
    s.advance_tracers(steps, xd, yd, L, L, periodic=periodic)
    got = s.download()
    s.close()
    hist = np.zeros_like(w, order="F")
    for _ in range(steps):
        w, hist = O.ab2csh_steps(w, hist, nx, nu, dt, 1)
        ux, uy = O.csh_velocity(w, nx, L, L)
        xp, yp = O.tracer_step(ux, uy, xp, yp, L, L, dt, periodic)
    assert rel_l2(got, w) < 1e-12
    assert np.all((xd > 0) & (xd <= L) & (yd > 0) & (yd <= L))
    ex, ey = _pdist(xd, xp, L).max(), _pdist(yd, yp, L).max()
    print(f"max position error x {ex:.3e} y {ey:.3e}", flush=True)
    assert ex < 1e-10 and ey < 1e-10


def test_tracers_steady_shear_closed_form():
    """omega = A cos y (one Fourier mode, nu = 0) is a steady Euler flow with
    ux = d psi/dy = -A sin y, uy = 0; particles on grid rows y0 = n dy see
    ux(y0) exactly (no interpolation across rows), so after K steps
    x = x0 - K dt A sin(y0) (mod Lx) and y = y0.  Particles whose path stays
    off the last column (x0 in [1, 4], |K dt A| < 0.7) move so in both
    addressing forms; the periodic form holds for particles anywhere."""
    nx, ny, a, dt, k = 64, 64, 1.3, 2e-3, 250
    print("tracers steady shear 64^2", flush=True)
    y = np.arange(ny) * L / ny
    w = O.r2c(np.asfortranarray(np.tile(a * np.cos(y), (nx, 1))))
    rng = np.random.default_rng(11)
    y0 = rng.integers(0, ny, 400) * (L / ny)
    x0 = np.concatenate([rng.uniform(1.0, 4.0, 200), rng.uniform(0.0, L, 200)])
    want = x0 - k * dt * a * np.sin(y0)
    for periodic in (True, False):
        xd, yd = np.ascontiguousarray(x0.copy()), np.ascontiguousarray(y0.copy())
        s = T.CshSolver(nx, ny, 0.0, dt, T.SCHEME_AB2)
        s.upload(w)
        s.advance_tracers(k, xd, yd, L, L, periodic=periodic)
        got = s.download()
        s.close()
        assert rel_l2(got, w) < 1e-14
        assert np.max(_pdist(yd, y0, L)) < 1e-13
        err = _pdist(xd, want, L)
        assert np.max(err[:200]) < 1e-11, periodic
        if periodic:
            assert np.max(err) < 1e-11


def test_tracers_reject_bad_input():
    s = T.CshSolver(32, 32, 1e-3, 1e-3)
    s.upload(O.ic2csh(32, 32)[1])
    xp = np.array([1.0, -0.5])
    yp = np.array([1.0, 1.0])
    with pytest.raises(T.TaraError):
        s.advance_tracers(1, xp, yp, L, L)
    assert xp[1] == -0.5
    s.close()
