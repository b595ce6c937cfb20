"""Pins the CPU oracle (oracle/tara_oracle.c) before it is trusted as the checker.

Each check is independent of the oracle's own formulas:
  * golden vectors written by the reference build (TARA Fortran + FFTW 3.3.9),
  * numpy's pocketfft (a library routine with the same axis order and DC/Nyquist
    handling), a long-double naive DFT,
  * closed forms and invariants of the equations.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

PI = np.pi


# ------------------------------------------------------------------ transforms
@pytest.mark.parametrize("n", [4, 8, 12, 16, 32])
def test_dft_vs_naive(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    for sign in (-1, 1):
        ref = np.array([np.sum(z * np.exp(sign * 2j * PI * j * k / n)) for k in range(n)])
        assert rel_l2(O.dft(z, sign), ref) < 1e-14


@pytest.mark.parametrize("nx,ny", [(16, 16), (64, 32), (32, 128)])
def test_r2c_c2r_2d_vs_numpy(nx, ny):
    rng = np.random.default_rng(nx * ny)
    x = np.asfortranarray(rng.standard_normal((nx, ny)))
    spec = O.r2c(x)
    assert rel_l2(spec, np.fft.rfft2(x.T).T) < 1e-14   # numpy: last axis halved -> transpose
    # non-Hermitian DC / Nyquist columns: c2r must ignore their imaginary parts
    # after the y pass, as numpy's irfft2 (complex y pass first, then c2r) does
    s = spec + 0.3 * (rng.standard_normal(spec.shape) + 1j * rng.standard_normal(spec.shape))
    s = np.asfortranarray(s)
    back = O.c2r(s, nx)
    assert rel_l2(back, np.fft.irfft2(s.T, s=(ny, nx)).T * nx * ny) < 1e-14


def test_r2c_c2r_3d_vs_numpy():
    rng = np.random.default_rng(3)
    x = np.asfortranarray(rng.standard_normal((16, 8, 32)))
    spec = O.r2c(x)
    assert rel_l2(spec, np.fft.rfftn(x.T).T) < 1e-14
    back = O.c2r(spec, 16)
    assert rel_l2(back / x.size, x) < 1e-14


# ------------------------------------------------------------------ derive2csh
@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("case,nu", [("ic", 0.01), ("rnd", 1e-3)])
def test_derive2csh_vs_reference(golden, n, case, nu):
    got = O.derive2csh(golden[f"d2csh_{n}_{case}_in"], n, nu)
    assert rel_l2(got, golden[f"d2csh_{n}_{case}_out"]) < 1e-13


def _grid(n):
    x = np.arange(n) * 2 * PI / n
    return np.meshgrid(x, x, indexing="ij")


def test_derive2csh_closed_form():
    """omega = cos x + cos 2y: psi = cos x + cos(2y)/4, u = (d_y psi, -d_x psi),
    u.grad omega = -1.5 sin x sin 2y exactly (band-limited, far below the cut)."""
    n = 64
    X, Y = _grid(n)
    w = O.r2c(np.asfortranarray(np.cos(X) + np.cos(2 * Y)))
    f = O.derive2csh(w, n, 0.0)
    fr = O.c2r(f, n) / n**2
    assert np.max(np.abs(fr - 1.5 * np.sin(X) * np.sin(2 * Y))) < 1e-12
    # viscous part: omega = cos(3x) has zero nonlinearity; f = -nu k^2 omega_hat
    w = O.r2c(np.asfortranarray(np.cos(3 * X)))
    f = O.derive2csh(w, n, 0.05)
    assert rel_l2(f, -0.05 * 9.0 * w) < 1e-12


def test_derive2csh_circular_mask_count():
    """With nu = 0 every mode at sqrt(kx^2+ky^2) >= (Nx+Ny)/6 is exactly zero:
    1374 of the 33*64 = 2112 half-plane modes at 64^2."""
    n = 64
    rng = np.random.default_rng(5)
    w = O.r2c(np.asfortranarray(rng.standard_normal((n, n))))
    f = O.derive2csh(w, n, 0.0)
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    cut = np.sqrt(kx**2 + ky**2) >= (2 * n / 2.0) / 3.0
    assert int(cut.sum()) == 1374
    assert np.all(f[cut] == 0)
    assert np.all(f[~cut] != 0)


def test_derive2csh_conserves_enstrophy():
    """For omega band-limited inside the dealias disc the truncated nonlinear term
    is the exact Galerkin projection: sum Re(conj(w) NL) = 0 and
    sum Re(conj(w) NL)/k^2 = 0 (enstrophy and energy), Hermitian weights."""
    n = 64
    rng = np.random.default_rng(9)
    w = O.r2c(np.asfortranarray(rng.standard_normal((n, n))))
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    k2 = (kx**2 + ky**2).astype(float)
    w[np.sqrt(k2) >= n / 3.0 - 1] = 0
    w = np.asfortranarray(w)
    nl = -O.derive2csh(w, n, 0.0)
    wt = np.where(kx == 0, 1.0, 2.0) * np.ones_like(k2)
    ens = np.sum(wt * np.real(np.conj(w) * nl))
    k2[0, 0] = 1.0
    en = np.sum(wt * np.real(np.conj(w) * nl) / k2)
    scale = np.sum(wt * np.abs(w) * np.abs(nl))
    assert abs(ens) / scale < 1e-13 and abs(en) / scale < 1e-13


# ------------------------------------------------------------------ updates
def test_ab2csh_vs_reference(golden):
    got = O.ab2csh(golden["ab2_in_w"], golden["ab2_in_fo"], golden["ab2_in_fn"], 64, 1e-3)
    assert rel_l2(got, golden["ab2_out"]) < 1e-15


@pytest.mark.parametrize("tag", ["asrun", "intent"])
def test_config1_trajectory_vs_reference(golden, tag):
    """2000 AB2 steps of the main2csh loop, 64^2: <= 1e-10 relative L2."""
    w, h = O.ab2csh_steps(golden[f"c1_{tag}_w0"], O.zeros_spec(64, 64), 64, 0.01, 1e-4, 1000)
    assert rel_l2(w, golden[f"c1_{tag}_w1000"]) < 1e-12
    w, h = O.ab2csh_steps(w, h, 64, 0.01, 1e-4, 1000)
    assert rel_l2(w, golden[f"c1_{tag}_w2000"]) < 1e-10


def test_ic_vs_reference(golden):
    omega, wk = O.ic2csh(64, 64, 2.0, 2.0, 25.0, 2.0, 3.0)
    assert np.max(np.abs(omega - golden["c1_asrun_omega0"])) == 0.0
    assert rel_l2(wk, golden["c1_asrun_w0"]) < 1e-15


def test_rk4_vs_reference(golden):
    w = O.rk4csh_steps(golden["rk4_w0"], 64, 0.01, 1e-4, 200)
    assert rel_l2(w, golden["rk4_w200"]) < 1e-12


def test_ab2_and_rk4_linear_decay_closed_form():
    """omega = cos(3x): the nonlinear term vanishes and each scheme reduces to a
    scalar recurrence of the decay rate lam = nu k^2 (k = 3)."""
    n, nu, dt, steps = 32, 0.1, 0.01, 50
    X, _ = _grid(n)
    w0 = O.r2c(np.asfortranarray(np.cos(3 * X)))
    lam = nu * 9.0
    # AB2 with zero history: y1 = y0 (1 - 1.5 lam dt); then y_{n+1} = y_n - dt lam (1.5 y_n - 0.5 y_{n-1})
    y_prev, y = 0.0, 1.0
    for _ in range(steps):
        f, fo = -lam * y, -lam * y_prev
        y_prev, y = y, y + (1.5 * f - 0.5 * fo) * dt
    w, _ = O.ab2csh_steps(w0, O.zeros_spec(n, n), n, nu, dt, steps)
    assert rel_l2(w, y * w0) < 1e-13
    # RK4 amplification factor R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, z = -lam dt
    z = -lam * dt
    amp = (1 + z + z * z / 2 + z**3 / 6 + z**4 / 24) ** steps
    w = O.rk4csh_steps(w0, n, nu, dt, steps)
    assert rel_l2(w, amp * w0) < 1e-13


def test_rk4_fourth_order():
    """RK4 over derive2csh converges at order 4 on the shear-layer flow."""
    n = 32
    _, w0 = O.ic2csh(n, n, 2.0, 2.0, 5.0, 2.0, 0.5)
    T = 0.2
    ref = O.rk4csh_steps(w0, n, 0.01, T / 64, 64)
    e1 = rel_l2(O.rk4csh_steps(w0, n, 0.01, T / 8, 8), ref)
    e2 = rel_l2(O.rk4csh_steps(w0, n, 0.01, T / 16, 16), ref)
    assert 12.0 < e1 / e2 < 20.0


# ------------------------------------------------------------------ derive3com
def test_derive3com_vs_reference(golden):
    st = [golden[f"d3com_in{k}"] for k in range(8)]
    L = 2 * PI
    got = O.derive3com(st, 16, L, L, L, 1.0 / 450.0, 1.0 / 450.0, 10.0)
    for k in (0, 1, 2, 3, 5, 6, 7):
        assert rel_l2(got[k], golden[f"d3com_out{k}"]) < 1e-13, k


def test_derive3com_uniform_state_is_steady():
    n = 8
    real = [np.ones((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)),
            np.ones((n, n, n)), 0.3 * np.ones((n, n, n)), -0.2 * np.ones((n, n, n)), 0.1 * np.ones((n, n, n))]
    st = [O.r2c(np.asfortranarray(r)) for r in real]
    L = 2 * PI
    f = O.derive3com(st, n, L, L, L, 1e-3, 1e-3, 10.0)
    for k in (0, 1, 2, 3, 5, 6, 7):
        assert np.max(np.abs(f[k])) == 0.0


def test_derive3com_induction_divergence_free(golden):
    """k . dB_hat/dt = 0: the induction term is a curl and diffusion keeps a
    solenoidal B solenoidal (B made divergence-free by projection first)."""
    st = [golden[f"d3com_in{k}"].copy() for k in range(8)]
    n = 16
    kx = np.arange(n // 2 + 1)[:, None, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :, None]
    kz = np.fft.fftfreq(n, 1.0 / n)[None, None, :]
    K = [kx + 0 * ky + 0 * kz, 0 * kx + ky + 0 * kz, 0 * kx + 0 * ky + kz]
    k2 = (K[0] ** 2 + K[1] ** 2 + K[2] ** 2).astype(float)
    k2[0, 0, 0] = 1.0
    kb = (K[0] * st[5] + K[1] * st[6] + K[2] * st[7]) / k2
    for c in range(3):
        st[5 + c] = np.asfortranarray(st[5 + c] - K[c] * kb)
    L = 2 * PI
    f = O.derive3com(st, n, L, L, L, 1e-3, 1e-3, 10.0)
    div = K[0] * f[5] + K[1] * f[6] + K[2] * f[7]
    scale = np.sqrt(sum(np.sum(np.abs(K[c] * f[5 + c]) ** 2) for c in range(3)))
    assert np.linalg.norm(div) / scale < 1e-13


# ------------------------------------------------------------------ derive2css
def _cos_field(n):
    x = np.arange(n) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    return X, Y, np.cos(X) + np.cos(2 * Y)


@pytest.mark.parametrize("lam,alpha,nu", [(10.0, 1.0, 0.0), (10.0, 0.7, 0.01), (2.5, 1.3, 0.1)])
def test_derive2css_closed_form(lam, alpha, nu):
    """omega = cos x + cos 2y: psi = cos x/(1+lam) + cos 2y/(4+lam), omega_d = k^2 psi,
    u . grad omega_d = -6 sin x sin 2y / ((1+lam)(4+lam)), so
    f = 6 alpha sin x sin 2y/((1+lam)(4+lam)) - nu cos x/(1+lam) - 16 nu cos 2y/(4+lam)."""
    n = 32
    X, Y, w = _cos_field(n)
    f = O.c2r(O.derive2css(O.r2c(w), n, alpha, nu, lam), n) / (n * n)
    exact = (6 * alpha * np.sin(X) * np.sin(2 * Y) / ((1 + lam) * (4 + lam))
             - nu * np.cos(X) / (1 + lam) - 16 * nu * np.cos(2 * Y) / (4 + lam))
    assert np.max(np.abs(f - exact)) < 1e-13


def test_derive2css_unscreened_is_reference_derive2csh(golden):
    """lambda = 0, alpha = 1 reduces derive2css to derive2csh; checked against the
    reference build's derive2csh vectors."""
    for n in (64, 128):
        w = golden[f"d2csh_{n}_rnd_in"]
        assert rel_l2(O.derive2css(w, n, 1.0, 1e-3, 0.0), golden[f"d2csh_{n}_rnd_out"]) < 1e-13


def test_derive2css_mask_count():
    """The circular dealias acts on the nonlinear part only: with nu = 0 the
    RHS vanishes on exactly the modes sqrt(kx^2+ky^2) >= (Nx+Ny)/6."""
    n = 64
    rng = np.random.default_rng(5)
    w = O.r2c(rng.standard_normal((n, n)))
    f = O.derive2css(w, n, 1.0, 0.0, 10.0)
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    cut = np.sqrt(kx * kx + ky * ky) >= (2 * n / 2.0) / 3.0
    assert int(cut.sum()) == 1374
    assert np.all(f[cut] == 0) and np.all(f[~cut][1:] != 0)


def test_rk4css_fourth_order():
    """Linear screened decay (single mode, nu > 0, no nonlinearity): the exact
    solution is w0 exp(-nu k^4/(k^2+lam) t); RK4's error falls ~16x per halving."""
    n, lam, nu = 16, 10.0, 0.5
    w = np.zeros((n // 2 + 1, n), np.complex128, order="F")
    w[3, 0] = 1.0 + 0.5j
    k2 = 9.0
    rate = nu * k2 * k2 / (k2 + lam)
    errs = []
    for steps in (10, 20, 40):
        got = O.rk4css_steps(w, n, 1.0, nu, lam, 0.4 / steps, steps)
        errs.append(abs(got[3, 0] - w[3, 0] * np.exp(-rate * 0.4)))
    assert 14 < errs[0] / errs[1] < 18 and 14 < errs[1] / errs[2] < 18


# ------------------------------------------------- 2D compressible families
COH = dict(pi=np.pi, lx=2 * np.pi, ly=2 * np.pi, mux=2e-4, muy=2e-4, cs=3.77)
COM = dict(pi=np.pi, lx=2 * np.pi, ly=2 * np.pi, mu=1e-3, eta=1e-4, cs=2.0)


@pytest.mark.parametrize("n", [32, 64])
def test_derive2coh_vs_reference(golden_fam2d, n):
    st = [golden_fam2d[f"coh_{n}_in{k}"] for k in range(3)]
    f = O.derive2coh(st, n, **COH)
    for k in range(3):
        assert rel_l2(f[k], golden_fam2d[f"coh_{n}_out{k}"]) < 1e-13


@pytest.mark.parametrize("n", [32, 64])
def test_derive2com_vs_reference(golden_fam2d, n):
    st = [golden_fam2d[f"com_{n}_in{k}"] for k in range(6)]
    f = O.derive2com(st, n, kf=0.0, a=0.0, **COM)
    for k in range(6):
        assert rel_l2(f[k], golden_fam2d[f"com_{n}_out{k}"]) < 1e-13


def test_derive2coh_uniform_state_is_steady():
    n = 32
    st = [O.r2c(np.ones((n, n))), O.r2c(np.zeros((n, n))), O.r2c(np.zeros((n, n)))]
    for f in O.derive2coh(st, n, **COH):
        assert np.max(np.abs(f)) < 1e-12


def test_derive2com_forcing_closed_form():
    """rho = 1, u = 0, B = 0, E uniform: every flux is uniform, so the momentum
    RHS is the (undealiased) forcing spectrum alone: F = A (-sin(kf y), sin(kf x))
    on the grid x(i) = (i-1) dx, y(j) = (j-1) dy; the other RHS vanish."""
    n, A = 32, 0.3
    x = np.arange(n) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    st = [O.r2c(a) for a in (np.ones((n, n)), np.zeros((n, n)), np.zeros((n, n)), 2 * np.ones((n, n)),
                             np.zeros((n, n)), np.zeros((n, n)))]
    f = O.derive2com(st, n, kf=1.0, a=A, **COM)
    assert rel_l2(f[1], O.r2c(-A * np.sin(Y))) < 1e-14
    assert rel_l2(f[2], O.r2c(A * np.sin(X))) < 1e-14
    for k in (0, 3, 4, 5):
        assert np.max(np.abs(f[k])) < 1e-12


def test_coh_steppers_closed_forms(golden_fam2d):
    """The reference's rk4_nk2coh and pc_nk2coh reduce to closed forms of the
    input slope (their derive2coh calls write unused locals): pinned bit for bit
    against the reference build."""
    s = golden_fam2d["pt_slope0"]
    w = golden_fam2d["coh_32_in0"]
    dt = 1e-3
    assert np.array_equal(golden_fam2d["pt_rk4nk"], w + (dt / 6.0) * (((s + 2.0 * s) + 2.0 * s) + s))
    assert np.array_equal(golden_fam2d["pt_pcnk"], w + ((s + s) * dt) / 2.0)


# ------------------------------------------------- counter-based initial conditions
PHILOX_KAT = [  # Philox-4x32-10 known-answer vectors published with the method
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert O.philox4x32_10(ctr, key) == want


def test_ic_random_properties():
    """Unit kinetic energy (from the real-space velocity), zero mean, Nyquist
    modes zero, and the spectrum a real field's (c2r then r2c returns it)."""
    n = 64
    wk = O.ic_random(n, n, 7, 4.0, 1.0)
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    k2 = np.where(kx * kx + ky * ky > 0, kx * kx + ky * ky, 1.0)
    ux = O.c2r(1j * ky * wk / k2, n) / n**2
    uy = O.c2r(-1j * kx * wk / k2, n) / n**2
    assert abs(0.5 * np.mean(ux * ux + uy * uy) - 1.0) < 1e-12
    assert wk[0, 0] == 0 and np.all(wk[n // 2, :] == 0) and np.all(wk[:, n // 2] == 0)
    assert rel_l2(O.r2c(O.c2r(wk, n)) / n**2, wk) < 1e-14
    assert not np.array_equal(wk, O.ic_random(n, n, 8, 4.0, 1.0))
