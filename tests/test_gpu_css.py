"""GPU parity of the screened-vorticity family (derive2css / ab2css / rk42css)
through the C ABI.  The reference routine reads uninitialised Lx, Ly and
lambda (src/derive2css.f95:11), so its own outputs pin nothing; the oracle's
restatement is pinned by the closed form, by the lambda = 0 reduction to the
reference's derive2csh and by RK4's order (tests/test_oracle.py)."""
import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rand_spec(nx, ny, seed):
    return O.r2c(np.asfortranarray(np.random.default_rng(seed).standard_normal((nx, ny))))


@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 256), (512, 512), (1024, 128)])
@pytest.mark.parametrize("alpha,lam", [(1.0, 10.0), (0.6, 2.5)])
def test_derive2css_vs_oracle(nx, ny, alpha, lam):
    print(f"derive2css {nx}x{ny} alpha={alpha} lam={lam}", flush=True)
    w = _rand_spec(nx, ny, nx + ny)
    before = w.copy()
    got = T.derive2css(nx, ny, nx // 2 + 1, alpha, 1e-3, w, lam)
    assert np.array_equal(w, before)
    assert rel_l2(got, O.derive2css(w, nx, alpha, 1e-3, lam)) < 1e-13


@pytest.mark.parametrize("n,nu", [(64, 0.01), (4096, 0.0)])
def test_derive2css_closed_form(n, nu):
    """nu = 0 at 4096^2: nu k^2 would lift the input transform's rounding noise
    at |k| ~ 2900 above the bound (as in the derive2csh closed-form test)."""
    print(f"derive2css closed form {n}", flush=True)
    lam, alpha = 10.0, 0.7
    x = np.arange(n) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    w = T.fft_r2c(np.cos(X) + np.cos(2 * Y))
    f = T.fft_c2r(T.derive2css(n, n, n // 2 + 1, alpha, nu, w, lam), n) / float(n * n)
    exact = (6 * alpha * np.sin(X) * np.sin(2 * Y) / ((1 + lam) * (4 + lam))
             - nu * np.cos(X) / (1 + lam) - 16 * nu * np.cos(2 * Y) / (4 + lam))
    assert np.max(np.abs(f - exact)) < 1e-11


def test_derive2css_unscreened_equals_derive2csh(golden):
    for n in (64, 128):
        w = golden[f"d2csh_{n}_rnd_in"]
        got = T.derive2css(n, n, n // 2 + 1, 1.0, 1e-3, w, 0.0)
        assert rel_l2(got, golden[f"d2csh_{n}_rnd_out"]) < 1e-13


def test_ab2css_matches_ab2csh(golden):
    a = T.ab2css(64, 64, 33, 1e-3, golden["ab2_in_w"], golden["ab2_in_fo"], golden["ab2_in_fn"])
    assert rel_l2(a, golden["ab2_out"]) < 1e-15


def test_rk42css_dropin_vs_oracle():
    n = 64
    w = _rand_spec(n, n, 3) * 1e-3
    got = T.rk42css(n, n, n // 2 + 1, 0.8, 0.01, 1e-3, w, 10.0)
    assert rel_l2(got, O.rk4css_steps(w, n, 0.8, 0.01, 10.0, 1e-3, 1)) < 1e-14


@pytest.mark.parametrize("scheme", [T.SCHEME_AB2, T.SCHEME_RK4])
def test_css_resident_trajectory_vs_oracle(scheme):
    """200 steps of the device-resident css solver against the oracle's loop
    (AB2 with zero initial history, or classic RK4)."""
    n, alpha, nu, lam, dt = 64, 1.0, 0.01, 10.0, 1e-3
    _, w0 = O.ic2csh(n, n)
    s = T.CshSolver(n, n, nu, dt, scheme, screened=True, alpha=alpha, lam=lam)
    s.upload(w0)
    s.advance(200)
    got = s.download()
    s.close()
    if scheme == T.SCHEME_RK4:
        ref = O.rk4css_steps(w0, n, alpha, nu, lam, dt, 200)
    else:
        ref = w0.copy(order="F")
        hist = np.zeros_like(ref, order="F")
        for _ in range(200):
            f = O.derive2css(ref, n, alpha, nu, lam)
            ref = O.ab2csh(ref, hist, f, n, dt)
            hist = f
    assert rel_l2(got, ref) < 1e-12
