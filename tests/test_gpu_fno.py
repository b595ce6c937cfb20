"""The FNO surrogate's device kernels (tara_fno_*) against the float64 CPU
oracle (oracle/fno_ref.py, torch FFTs + autograd) on the same parameters and
inputs, and the ensemble's device data paths that feed it.

Tolerances (relative L2): FP32 precision, forward 1e-5 and each gradient block
1e-5 (FP32 MFMA accumulation against float64; measured about 2e-7); bfloat16
precision, forward 2e-2 and each gradient block 2e-2 (bfloat16 operands, FP32
accumulation; measured about 3e-3)."""
import numpy as np
import pytest

import tara_amd as T
from oracle import fno_ref as F

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _fields(n, nb, seed):
    """Smooth random fields with a decaying spectrum and a slightly shifted target."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    amp = np.exp(-(kk / 4.0) ** 2) * (kk > 0)
    out = []
    for _ in range(2 * nb):
        ph = np.exp(2j * np.pi * rng.random((n, n)))
        out.append(np.real(np.fft.ifft2(amp * ph)) * n)
    om = np.array(out[:nb], np.float32)
    tgt = (0.9 * om + 0.2 * np.array(out[nb:], np.float32)).astype(np.float32)
    return om, tgt


def _blocks(cfg, g):
    o, out = 0, {}
    for name, shape in cfg.sizes():
        k = int(np.prod(shape))
        out[name] = g[o:o + k]
        o += k
    return out


@pytest.mark.parametrize("prec", [T.FNO_FP32, T.FNO_BF16])
@pytest.mark.parametrize("n,width,mx,my,layers,nb", [(32, 8, 6, 8, 2, 3), (64, 16, 8, 16, 3, 2)])
def test_fno_forward_and_grads_vs_oracle(prec, n, width, mx, my, layers, nb):
    print(f"fno prec={prec} n={n} width={width} modes={mx}x{my} layers={layers} batch={nb}", flush=True)
    cfg = F.Config(n, width, mx, my, layers)
    net = T.FNO(n, width, mx, my, layers, max_batch=nb, precision=prec, seed=11)
    assert net.nparams == cfg.nparams()
    rng = np.random.default_rng(n + width)
    p = net.get_params()
    # a trained-looking network: larger projection than the near-identity start
    b = _blocks(cfg, p)
    b["w2"][:] = rng.standard_normal(b["w2"].size) * 0.3
    b["b2"][:] = 0.05
    net.set_params(p)
    om, tgt = _fields(n, nb, seed=n)
    d_om = torch.from_numpy(om).cuda()
    d_tgt = torch.from_numpy(tgt).cuda()
    d_out = torch.empty_like(d_om)
    net.forward(nb, d_om, d_out)
    pred, lref, gref = F.run(cfg, p.astype(np.float64), om, tgt)
    got = d_out.cpu().numpy()
    ftol, gtol = (1e-5, 1e-5) if prec == T.FNO_FP32 else (2e-2, 2e-2)
    assert _rel(got, pred) < ftol
    # the increment out - omega alone, which is what the layers compute
    assert _rel(got - om, pred - om) < (1e-4 if prec == T.FNO_FP32 else 3e-2)
    loss, g = net.grads(nb, d_om, d_tgt)
    assert abs(loss - lref) < 10 * ftol * max(lref, 1e-3)
    gb, rb = _blocks(cfg, g), _blocks(cfg, gref)
    for name, _ in cfg.sizes():
        if rb[name].size < 8 and prec == T.FNO_BF16:
            continue
        e = _rel(gb[name], rb[name])
        print(f"  grad {name}: rel {e:.2e}", flush=True)
        assert e < gtol, (name, e)


def test_fno_production_shape_bf16_vs_oracle():
    n, width, mx, my, layers, nb = 256, 64, 16, 32, 4, 2
    print(f"fno production shape n={n} width={width} modes={mx}x{my} batch={nb}", flush=True)
    cfg = F.Config(n, width, mx, my, layers)
    net = T.FNO(n, width, mx, my, layers, max_batch=nb, precision=T.FNO_BF16, seed=2)
    p = net.get_params()
    b = _blocks(cfg, p)
    b["w2"][:] = np.random.default_rng(0).standard_normal(b["w2"].size) * 0.3
    net.set_params(p)
    om, tgt = _fields(n, nb, seed=7)
    d_om, d_tgt = torch.from_numpy(om).cuda(), torch.from_numpy(tgt).cuda()
    d_out = torch.empty_like(d_om)
    net.forward(nb, d_om, d_out)
    pred, lref, gref = F.run(cfg, p.astype(np.float64), om, tgt)
    assert _rel(d_out.cpu().numpy() - om, pred - om) < 3e-2
    loss, g = net.grads(nb, d_om, d_tgt)
    gb, rb = _blocks(cfg, g), _blocks(cfg, gref)
    for name in ("r", "w", "w1", "w2"):
        e = _rel(gb[name], rb[name])
        print(f"  grad {name}: rel {e:.2e}", flush=True)
        assert e < 2e-2, (name, e)


def test_fno_training_reduces_loss():
    n, nb = 32, 4
    net = T.FNO(n, 8, 6, 8, 2, max_batch=nb, precision=T.FNO_BF16, seed=5)
    om, tgt = _fields(n, nb, seed=1)
    d_om, d_tgt = torch.from_numpy(om).cuda(), torch.from_numpy(tgt).cuda()
    losses = [net.train_step(nb, d_om, d_tgt, 3e-3) for _ in range(60)]
    print(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]


def test_fno_rel_err():
    n, nb = 32, 3
    net = T.FNO(n, 8, 6, 8, 1, max_batch=nb, seed=1)
    om, tgt = _fields(n, nb, seed=3)
    d_a, d_r = torch.from_numpy(om).cuda(), torch.from_numpy(tgt).cuda()
    e = net.rel_err(nb, d_a, d_r)
    want = [np.linalg.norm(om[b] - tgt[b]) / np.linalg.norm(tgt[b]) for b in range(nb)]
    assert np.allclose(e, want, rtol=1e-5)


def test_fno_gather_pairs():
    """The training batch gather equals indexing, and an index outside the
    pair store is refused before anything runs."""
    n, nb, npairs = 32, 8, 50
    net = T.FNO(n, 8, 6, 8, 1, max_batch=nb, seed=2)
    rng = np.random.default_rng(4)
    snaps = torch.from_numpy(rng.standard_normal((npairs + 1, n, n)).astype(np.float32)).cuda()
    inp, tgt = snaps[:-1], snaps[1:]
    xb = torch.empty((nb, n, n), dtype=torch.float32, device="cuda")
    yb = torch.empty_like(xb)
    idx = rng.permutation(npairs)[:nb]
    net.gather_pairs(idx, inp, tgt, xb, yb)
    torch.cuda.synchronize()
    assert torch.equal(xb.cpu(), inp.cpu()[idx]) and torch.equal(yb.cpu(), tgt.cpu()[idx])
    with pytest.raises(T.TaraError):
        net.gather_pairs(np.array([0, npairs]), inp, tgt, xb, yb)


def test_fno_rejects_bad_shapes():
    with pytest.raises(T.TaraError):
        T.FNO(30, 8, 6, 8, 1)  # n not a multiple of 8
    with pytest.raises(T.TaraError):
        T.FNO(32, 8, 17, 8, 1)  # 2 mx > n


def test_ensemble_real_space_paths():
    n, nu, dt = 64, 1e-3, 5e-4
    e = T.Ensemble([n, n], [nu, nu], dt)
    s = T.CshSolver(n, n, nu=nu, dt=dt, scheme=T.SCHEME_RK4)
    e.ic_random(1, seed=9, k0=5.0, energy=1.0)
    s.ic_random(seed=9, k0=5.0, energy=1.0)
    w_ens, w_one = e.download(1), s.download()
    assert np.array_equal(w_ens, w_one)
    # real space out: float32 of the FP64 inverse transform
    buf = torch.zeros((n, n), dtype=torch.float32).cuda()
    e.get_omega(1, buf)
    e.sync()
    om_ens = buf.cpu().numpy()
    om_ref = np.fft.irfft2(np.asarray(w_one).T, s=(n, n))  # [ky][kx] -> [y][x]
    assert _rel(om_ens, om_ref) < 1e-6
    # real space in: the run's spectrum is the FP64 forward transform of the float field
    e.set_omega(0, buf)
    e.sync()
    w_in = e.download(0)
    want = np.fft.rfft2(om_ens.astype(np.float64)).T
    assert _rel(w_in, want) < 1e-12
    # and both runs advance identically from (numerically) the same state
    e.advance(20)
    e.sync()
    s.advance(20)
    assert _rel(e.download(1), s.download()) < 1e-12
    e.close()
    s.close()
