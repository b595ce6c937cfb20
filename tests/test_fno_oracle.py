"""The FNO surrogate's CPU oracle (oracle/fno_ref.py), pinned by what the
mathematics fixes: the spectral convolution of a plane wave is the plane wave
times its mode's weight (zero outside the retained modes), the parameter layout
has the documented size, and autograd's gradient of the loss agrees with
central finite differences."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import fno_ref as F  # noqa: E402


def test_param_count_production_shape():
    # 4 layers, width 64, kx < 16, -16 <= ky < 16: 4 * 512 * 2 * 64 * 64 spectral weights
    cfg = F.Config(256, 64, 16, 32, 4)
    spectral = 4 * 512 * 2 * 64 * 64
    rest = 64 + 64 + 4 * 64 * 64 + 4 * 64 + 128 * 64 + 128 + 128 + 1
    assert cfg.nparams() == spectral + rest


@pytest.mark.parametrize("kx,ky,kept", [(2, -1, True), (0, 1, True), (3, -2, True), (5, 0, False), (1, 2, False)])
def test_spectral_conv_plane_wave(kx, ky, kept):
    n, mx, my = 16, 4, 4
    cfg = F.Config(n, 1, mx, my, 1)
    rng = np.random.default_rng(kx * 10 + ky + 50)
    r = torch.tensor(rng.standard_normal((mx * my, 2, 1, 1)))
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    th = 2 * np.pi * (kx * x + ky * y) / n
    v = torch.tensor(np.cos(th))[None, None]
    out = F.spectral_conv(cfg, v, r).numpy()[0, 0]
    if not kept:
        assert np.abs(out).max() < 1e-12
        return
    row = ky if ky >= 0 else ky + my
    m = row * mx + kx
    rr, ri = float(r[m, 0, 0, 0]), float(r[m, 1, 0, 0])
    if kx == 0:
        # cos(ky y) holds +ky and -ky: both rows are retained here, each with its own weight
        m2 = ((-ky) % my) * mx
        r2r, r2i = float(r[m2, 0, 0, 0]), float(r[m2, 1, 0, 0])
        want = 0.5 * (rr * np.cos(th) - ri * np.sin(th)) + 0.5 * (r2r * np.cos(th) + r2i * np.sin(th))
    else:
        want = rr * np.cos(th) - ri * np.sin(th)
    assert np.abs(out - want).max() < 1e-12


def test_gradient_matches_finite_differences():
    cfg = F.Config(8, 3, 2, 4, 2)
    rng = np.random.default_rng(3)
    p = rng.standard_normal(cfg.nparams()) * 0.3
    om = rng.standard_normal((2, 8, 8))
    tgt = om + 0.3 * rng.standard_normal((2, 8, 8))
    _, l0, g = F.run(cfg, p, om, tgt)
    idx = rng.choice(cfg.nparams(), 25, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (F.run(cfg, pp, om, tgt)[1] - F.run(cfg, pm, om, tgt)[1]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_zero_projection_is_identity():
    cfg = F.Config(16, 4, 3, 4, 2)
    rng = np.random.default_rng(5)
    flat = torch.tensor(rng.standard_normal(cfg.nparams()))
    blocks = F.unflatten(cfg, flat)  # views into flat
    blocks["w2"][:] = 0
    blocks["b2"][:] = 0
    p = flat.numpy()
    om = rng.standard_normal((3, 16, 16))
    assert np.abs(F.run(cfg, p, om) - om).max() < 1e-14
