"""Device tracer particles (src/main2csh.f95:187-265) through the C ABI:
against the reference program's own text snapshots (fixture written by
tests/golden/make_golden.py snap), against the oracle's restatement step by
step over a run of the driver's shear-layer flow, and against the closed form
of a steady shear flow."""
import os

import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import ROOT, rel_l2

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


def _golden_snap():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_snap.npz"))


def test_tracers_and_fields_vs_reference_program():
    """The reference program's config-1 run (64^2 shear layer, nu = 0.01,
    dt = 1e-3, AB2, 256 particles) against the device: omega, ux, uy and the
    positions after 1001 steps (one advance_tracers call), and after 1 step.
    Particles that enter the last row are left out (the program reads past
    its 32 KiB arrays into the heap there; the fixture marks them)."""
    g = _golden_snap()
    nx, nu, dt = 64, 0.01, 1e-3
    print("tracers vs reference program 64^2", flush=True)
    for steps, t in ((1, 0), (1001, 1000)):
        xd = np.ascontiguousarray(g["s64_start"][:, 0])
        yd = np.ascontiguousarray(g["s64_start"][:, 1])
        s = T.CshSolver(nx, nx, nu, dt, T.SCHEME_AB2)
        s.ic_shear(2.0, 2.0, 25.0, 2.0, 0.0234375)
        s.advance_tracers(steps, xd, yd, L, L)
        om, ux, uy = s.fields()
        s.close()
        f = g[f"s64_fields_{t}"]
        for col, got in ((2, om), (3, ux), (4, uy)):
            ref = f[:, col].reshape(nx, nx)
            assert np.max(np.abs(got - ref)) < 1e-10 * np.max(np.abs(ref)), (t, col)
        p = g[f"s64_particles_{t}"]
        ok = ~g["s64_last_row"]
        err = max(np.max(np.abs(p[ok, 0] - xd[ok])), np.max(np.abs(p[ok, 1] - yd[ok])))
        assert err < 1e-10, (t, err)


def test_tracers_past_the_end_reads_vs_reference_program():
    """One step of 2048 particles at 128^2 against the reference program,
    every particle (last row, last column and corner cell included: the
    program's arrays are mmap-allocated there, so its reads past their end
    give 0, as the device's driver addressing)."""
    g = _golden_snap()
    nx = 128
    xd = np.ascontiguousarray(g["s128_start"][:, 0])
    yd = np.ascontiguousarray(g["s128_start"][:, 1])
    s = T.CshSolver(nx, nx, 0.01, 1e-3, T.SCHEME_AB2)
    s.ic_shear(2.0, 2.0, 25.0, 2.0, 0.0234375)
    s.advance_tracers(1, xd, yd, L, L)
    s.close()
    p = g["s128_particles_0"]
    assert np.max(np.abs(p[:, 0] - xd)) < 1e-12 and np.max(np.abs(p[:, 1] - yd)) < 1e-12


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
