"""GPU parity of the 2D compressible families (derive2coh / derive2com and
their steppers) through the C ABI, against the reference build's vectors
(tests/golden/golden_fam2d.npz) and the oracle's restatement.

Bars: single RHS calls 1e-12 relative L2 (two FFT factorisations, and the
real-space sums formed before one transform instead of after nine); 200-step
AB2 trajectories 1e-10; pointwise updates 1e-15."""
import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import rel_l2

pytestmark = pytest.mark.gpu

PI = np.pi
COH = dict(pi=PI, lx=2 * PI, ly=2 * PI, mux=2e-4, muy=2e-4, cs=3.77)
COM = dict(pi=PI, lx=2 * PI, ly=2 * PI, mu=1e-3, eta=1e-4, cs=2.0)


def coh_params(dt=0.0):
    return T.CohParams(COH["pi"], COH["lx"], COH["ly"], COH["mux"], COH["muy"], COH["cs"], dt)


def com_params(kf=0.0, a=0.0, dt=0.0):
    return T.ComParams(COM["pi"], COM["lx"], COM["ly"], COM["mu"], COM["eta"], COM["cs"], kf, a, dt)


@pytest.mark.parametrize("n", [32, 64])
def test_derive2coh_vs_reference(golden_fam2d, n):
    print(f"derive2coh {n}", flush=True)
    st = [golden_fam2d[f"coh_{n}_in{k}"] for k in range(3)]
    f = T.derive2coh(n, n, n // 2 + 1, coh_params(), st)
    for k in range(3):
        assert rel_l2(f[k], golden_fam2d[f"coh_{n}_out{k}"]) < 1e-12


@pytest.mark.parametrize("n", [32, 64])
def test_derive2com_vs_reference(golden_fam2d, n):
    print(f"derive2com {n}", flush=True)
    st = [golden_fam2d[f"com_{n}_in{k}"] for k in range(6)]
    f = T.derive2com(n, n, n // 2 + 1, com_params(), st)
    for k in range(6):
        assert rel_l2(f[k], golden_fam2d[f"com_{n}_out{k}"]) < 1e-12


def _smooth_state(n, seed, nf):
    rng = np.random.default_rng(seed)
    x = np.arange(n) * 2 * PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")

    def smooth():
        f = np.zeros((n, n))
        for _ in range(6):
            k = rng.integers(-4, 5, size=2)
            f += rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + rng.uniform(0, 2 * PI))
        return f

    reals = [1.0 + 0.05 * smooth()] + [0.2 * smooth() for _ in range(nf - 1)]
    if nf == 6:
        reals[3] = 2.0 + 0.1 * smooth()
    return [O.r2c(np.asfortranarray(a)) for a in reals]


@pytest.mark.parametrize("nx,ny", [(16, 16), (128, 64), (256, 256), (64, 256)])
def test_derive2coh_vs_oracle(nx, ny):
    print(f"derive2coh oracle {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx + ny)
    dens = 1.0 + 0.02 * rng.standard_normal((nx, ny))
    st = [O.r2c(np.asfortranarray(a)) for a in (dens, 0.1 * rng.standard_normal((nx, ny)),
                                                 0.1 * rng.standard_normal((nx, ny)))]
    f = T.derive2coh(nx, ny, nx // 2 + 1, coh_params(), st)
    ref = O.derive2coh(st, nx, **COH)
    for k in range(3):
        assert rel_l2(f[k], ref[k]) < 1e-12


@pytest.mark.parametrize("n", [16, 128, 256])
def test_derive2com_with_forcing_vs_oracle(n):
    print(f"derive2com forcing {n}", flush=True)
    st = _smooth_state(n, 3 * n, 6)
    f = T.derive2com(n, n, n // 2 + 1, com_params(kf=1.0, a=0.3), st)
    ref = O.derive2com(st, n, kf=1.0, a=0.3, **COM)
    for k in range(6):
        assert rel_l2(f[k], ref[k]) < 1e-12


def test_derive2com_forcing_closed_form():
    n, A = 64, 0.3
    x = np.arange(n) * 2 * PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    st = [T.fft_r2c(a) for a in (np.ones((n, n)), np.zeros((n, n)), np.zeros((n, n)), 2 * np.ones((n, n)),
                                 np.zeros((n, n)), np.zeros((n, n)))]
    f = T.derive2com(n, n, n // 2 + 1, com_params(kf=1.0, a=A), st)
    assert rel_l2(f[1], O.r2c(-A * np.sin(Y))) < 1e-14
    assert rel_l2(f[2], O.r2c(A * np.sin(X))) < 1e-14
    for k in (0, 3, 4, 5):
        assert np.max(np.abs(f[k])) < 1e-12


def test_coh_resident_ab2_vs_reference(golden_fam2d):
    n = 32
    s = T.F2Solver(n, n, coh_params(dt=1e-4))
    s.upload([golden_fam2d[f"coh_{n}_in{k}"] for k in range(3)])
    s.advance(200)
    got = s.download()
    s.close()
    for k in range(3):
        assert rel_l2(got[k], golden_fam2d[f"cohtraj_{n}_w{k}"]) < 1e-10


def test_com_resident_ab2_vs_reference(golden_fam2d):
    n = 32
    s = T.F2Solver(n, n, com_params(dt=1e-4))
    s.upload([golden_fam2d[f"com_{n}_in{k}"] for k in range(6)])
    s.advance(200)
    got = s.download()
    s.close()
    for k in range(6):
        assert rel_l2(got[k], golden_fam2d[f"comtraj_{n}_w{k}"]) < 1e-10


def test_coh_pointwise_steppers_vs_reference(golden_fam2d):
    n, nh, dt = 32, 17, 1e-3
    st = [golden_fam2d[f"coh_32_in{k}"] for k in range(3)]
    sl = [golden_fam2d[f"pt_slope{k}"] for k in range(3)]
    fo = [golden_fam2d[f"pt_fold{k}"] for k in range(3)]
    for k, a in enumerate(T.ab2coh(n, n, nh, dt, st, fo, sl)):
        assert rel_l2(a, golden_fam2d[f"pt_ab2coh{k}"]) < 1e-15
    for k, a in enumerate(T.euler2coh(n, n, nh, dt, st, sl)):
        assert rel_l2(a, golden_fam2d[f"pt_euler{k}"]) < 1e-15
    assert rel_l2(T.rk4coh_field(n, n, nh, dt, st[0], sl[0]), golden_fam2d["pt_rk4nk"]) < 1e-15
    assert rel_l2(T.pccoh_field(n, n, nh, dt, st[0], sl[0]), golden_fam2d["pt_pcnk"]) < 1e-15


def test_ab2com_vs_oracle_formula(golden_fam2d):
    n, nh, dt = 32, 17, 1e-3
    st = [golden_fam2d[f"com_32_in{k}"] for k in range(6)]
    fo = [golden_fam2d[f"com_32_out{k}"] for k in range(6)]
    fn = [a * 0.5 for a in fo]
    got = T.ab2com(n, n, nh, dt, st, fo, fn)
    for k in range(6):
        assert rel_l2(got[k], st[k] + (1.5 * fn[k] - 0.5 * fo[k]) * dt) < 1e-15
