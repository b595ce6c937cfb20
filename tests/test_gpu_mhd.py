"""GPU parity of the 3D MHD path (derive3com / ab3com), through the C ABI."""
import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import rel_l2

pytestmark = pytest.mark.gpu
PI = 3.14159265358979323846
L = 2.0 * PI  # input.ini Lx = 2.0, multiplied by pi (src/main3com.f95:203)


def smooth_state(n, seed, nmodes=6, kmax=3):
    rng = np.random.default_rng(seed)
    x = np.arange(n) * L / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")

    def field():
        f = np.zeros((n, n, n))
        for _ in range(nmodes):
            k = rng.integers(-kmax, kmax + 1, size=3)
            f += rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 2 * PI))
        return f

    rho = 1.0 + 0.1 * field() / nmodes
    u = [0.1 * field() for _ in range(3)]
    B = [0.1 * field() for _ in range(3)]
    reals = [rho, rho * u[0], rho * u[1], rho * u[2], np.ones_like(rho), B[0], B[1], B[2]]
    return [T.fft_r2c(np.asfortranarray(r)) for r in reals]


def params(mu, eta, dt=0.0, cs=10.0):
    return T.MhdParams(L, L, L, mu, eta, cs, dt)


LIVE = (0, 1, 2, 3, 5, 6, 7)


def test_derive3com_vs_reference(golden):
    st = [golden[f"d3com_in{k}"] for k in range(8)]
    got = T.derive3com(16, 16, 16, 9, params(1.0 / 450.0, 1.0 / 450.0), st)
    for k in LIVE:
        assert rel_l2(got[k], golden[f"d3com_out{k}"]) < 1e-13, k
    assert np.all(got[4] == 0)


@pytest.mark.parametrize("n", [32, 64])
def test_derive3com_vs_oracle(n):
    print(f"derive3com {n}^3", flush=True)
    st = smooth_state(n, seed=n)
    got = T.derive3com(n, n, n, n // 2 + 1, params(1e-3, 2e-3), st)
    ref = O.derive3com(st, n, L, L, L, 1e-3, 2e-3, 10.0)
    for k in LIVE:
        assert rel_l2(got[k], ref[k]) < 1e-12, k


def test_derive3com_uniform_state_is_steady():
    n = 32
    real = [np.ones((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)),
            np.ones((n, n, n)), 0.3 * np.ones((n, n, n)), -0.2 * np.ones((n, n, n)), 0.1 * np.ones((n, n, n))]
    st = [T.fft_r2c(np.asfortranarray(r)) for r in real]
    f = T.derive3com(n, n, n, n // 2 + 1, params(1e-3, 1e-3), st)
    for k in LIVE:
        assert np.max(np.abs(f[k])) == 0.0


def test_ab3com_vs_oracle():
    n = 16
    st, fo, fn = smooth_state(n, 1), smooth_state(n, 2), smooth_state(n, 3)
    got = T.ab3com(n, n, n, 9, 1e-3, st, fo, fn)
    ref = O.ab3com(st, fo, fn, n, 1e-3)
    for k in range(8):
        assert rel_l2(got[k], ref[k]) < 1e-15
    assert np.array_equal(got[4], st[4])


@pytest.mark.parametrize("n,steps", [(32, 6), (64, 2)])
def test_mhd_resident_ab2_vs_oracle(n, steps):
    print(f"mhd ab2 {n}^3 x {steps}", flush=True)
    st = smooth_state(n, seed=100 + n)
    dt, mu, eta = 5e-4, 1e-3, 1e-3
    s = T.MhdSolver(n, n, n, params(mu, eta, dt))
    s.upload(st)
    s.advance(steps)
    got = s.download()
    s.close()
    ref = [a.copy(order="F") for a in st]
    hist = [np.zeros_like(a, order="F") for a in st]
    for _ in range(steps):
        f = O.derive3com(ref, n, L, L, L, mu, eta, 10.0)
        ref = O.ab3com(ref, hist, f, n, dt)
        hist = f
    for k in range(8):
        assert rel_l2(got[k], ref[k]) < 1e-12, k


def test_mhd_256_one_step_consistency():
    """At 256^3 (the benchmark size): one device-resident AB2 step equals
    drop-in derive3com + ab3com with zero history, and the induction RHS of a
    solenoidal B stays solenoidal (k . dB/dt = 0)."""
    n = 256
    st = smooth_state(n, seed=5, nmodes=4)
    kx = np.arange(n // 2 + 1)[:, None, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :, None]
    kz = np.fft.fftfreq(n, 1.0 / n)[None, None, :]
    K = [kx + 0 * ky + 0 * kz, 0 * kx + ky + 0 * kz, 0 * kx + 0 * ky + kz]
    k2 = (K[0] ** 2 + K[1] ** 2 + K[2] ** 2).astype(float)
    k2[0, 0, 0] = 1.0
    kb = (K[0] * st[5] + K[1] * st[6] + K[2] * st[7]) / k2
    for c in range(3):
        st[5 + c] = np.asfortranarray(st[5 + c] - K[c] * kb)
    del kb, k2
    prm = params(1e-3, 1e-3, 5e-4)
    f = T.derive3com(n, n, n, n // 2 + 1, prm, st)
    div = K[0] * f[5] + K[1] * f[6] + K[2] * f[7]
    scale = np.sqrt(sum(np.sum(np.abs(K[c] * f[5 + c]) ** 2) for c in range(3)))
    assert np.linalg.norm(div) / scale < 1e-12
    del div
    zero = [np.zeros_like(st[0], order="F") for _ in range(8)]
    ref = T.ab3com(n, n, n, n // 2 + 1, 5e-4, st, zero, f)
    s = T.MhdSolver(n, n, n, prm)
    s.upload(st)
    s.advance(1)
    got = s.download()
    s.close()
    for k in range(8):
        assert rel_l2(got[k], ref[k]) < 1e-14, k


@pytest.mark.parametrize("n", [16, 64])
def test_mhd_ic_abc_vs_numpy(n):
    """The device ABC-flow initial state (src/main3com.f95:290-324) against numpy's
    FFT of the same real fields, and a uniform-B, divergence-free start."""
    lx = 2.0 * np.pi
    rho0, u0, b0, p0 = 1.0, 0.1, 0.1, 1.0
    s = T.MhdSolver(n, n, n, params(1.0 / 450, 1.0 / 450))
    s.ic_abc(rho0, u0, b0, p0)
    got = s.download()
    s.close()
    x = np.arange(n) * lx / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")  # (i, j, k) = (x, y, z), Fortran order below
    ux = u0 * (np.sin(Z) + np.cos(Y))
    uy = u0 * (np.sin(X) + np.cos(Z))
    uz = u0 * (np.sin(Y) + np.cos(X))
    E = p0 + 0.5 * rho0 * (ux ** 2 + uy ** 2 + uz ** 2) + 1.5 * b0 ** 2
    fields = [rho0 + 0 * X, rho0 * ux, rho0 * uy, rho0 * uz, E, b0 + 0 * X, b0 + 0 * X, b0 + 0 * X]
    for f, r in enumerate(fields):
        # reference layout: (Nh, Ny, Nz) half spectrum over the first (x) index
        want = np.fft.rfftn(r.transpose(2, 1, 0), axes=(0, 1, 2)).transpose(2, 1, 0)
        assert rel_l2(got[f], want) < 1e-13, f


def test_mhd_energy_closed_form_abc():
    """The device diagnostics of the ABC-flow state (rho = 1, u = u0 (A sin z +
    C cos y, B sin x + A cos z, C sin y + B cos x), uniform B = b0 on each
    axis, E = p0 + u^2/2 + B^2/2) against their closed forms: <CS^2 rho> =
    CS^2, <|u|^2>/2 = u0^2 (A^2+B^2+C^2)/2, <E> = p0 + that + 3 b0^2/2,
    <|B|^2>/2 = 3 b0^2/2, A = curl B/k^2 = 0 (uniform B), div B = 0; the
    flow is Beltrami (curl u = u), and the driver's omega has its y component
    negated (src/main3com.f95:890), so |omega| = |u| pointwise: <|omega|^2> =
    u0^2 (A^2+B^2+C^2), u.omega = ux^2 - uy^2 + uz^2, whose mean is
    u0^2 (A^2 + C^2 - B^2 - A^2 + C^2 + B^2)/2 = u0^2 C^2; HT = Hf (A = 0,
    and <u.B> = <omega.B> = 0 for a uniform B and zero-mean u)."""
    n, u0, b0, p0 = 32, 0.1, 0.1, 1.0
    a, b, c = 1.0, 0.7, 0.4
    s = T.MhdSolver(n, n, n, params(1.0 / 450, 1.0 / 450))
    s.ic_abc(1.0, u0, b0, p0, a, b, c, 1.0)
    got = s.energy()
    s.close()
    ke = u0 ** 2 * (a * a + b * b + c * c) / 2
    want = {0: 100.0, 1: ke, 2: p0 + ke + 1.5 * b0 ** 2, 3: 1.5 * b0 ** 2, 5: 2 * ke, 6: u0 ** 2 * c * c,
            7: 0.0, 8: u0 ** 2 * c * c, 9: 0.0}
    for k, v in want.items():
        assert abs(got[k] - v) < 1e-13 * max(1.0, abs(v)), (k, got[k], v)
    # <|omega|> = <|u|>: the mean speed of the ABC field, by quadrature on a finer grid
    m = 96
    x = np.arange(m) * 2 * np.pi / m
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij", sparse=True)
    speed = u0 * np.sqrt((a * np.sin(Z) + c * np.cos(Y)) ** 2 + (b * np.sin(X) + a * np.cos(Z)) ** 2
                         + (c * np.sin(Y) + b * np.cos(X)) ** 2)
    assert abs(got[4] - speed.mean()) < 1e-3 * got[4]
