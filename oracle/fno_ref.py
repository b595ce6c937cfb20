"""TEST INFRASTRUCTURE ONLY -- CPU oracle of the Fourier-neural-operator surrogate.

A float64 restatement, on torch's CPU FFTs and autograd, of the network that
``tara_amd/csrc/fno.hip`` runs on the matrix cores (the surrogate has no
counterpart in the reference; its definition is the one in that file's
header).  It is written the textbook way -- rfft2, keep the retained corner
blocks, multiply per mode by a complex matrix, irfft2 -- rather than as the
truncated-DFT GEMMs of the device code, so that the two share no derivation.
Gradients come from autograd on the same loss (batch mean of relative L2
errors).

Only tests/ import this module.
"""
import numpy as np
import torch


class Config:
    def __init__(self, n, width, mx, my, layers):
        self.n, self.c, self.mx, self.my, self.layers = n, width, mx, my, layers
        self.h = 2 * width

    @property
    def modes(self):
        return self.mx * self.my

    def sizes(self):
        """(name, shape) of each parameter block, in the flat vector's order."""
        c, h, m, l = self.c, self.h, self.modes, self.layers
        return [("wl", (c,)), ("bl", (c,)), ("r", (l, m, 2, c, c)), ("w", (l, c, c)), ("b", (l, c)),
                ("w1", (h, c)), ("b1", (h,)), ("w2", (h,)), ("b2", (1,))]

    def nparams(self):
        return sum(int(np.prod(s)) for _, s in self.sizes())


def unflatten(cfg, flat):
    """Split a flat parameter tensor into named blocks (views)."""
    out, o = {}, 0
    for name, shape in cfg.sizes():
        k = int(np.prod(shape))
        out[name] = flat[o:o + k].reshape(shape)
        o += k
    return out


def spectral_conv(cfg, v, r):
    """K v for one layer: v (B, C, N, N) real, r (MX*MY, 2, C, C) with mode
    index ky_r*MX + kx and row ky_r holding ky = ky_r (ky_r < MY/2) or
    ky_r - MY; [.., 0] real and [.., 1] imaginary parts, [in][out]."""
    n, mx, my = cfg.n, cfg.mx, cfg.my
    xf = torch.fft.rfft2(v)                                    # (B, C, N, N/2+1)
    w = torch.complex(r[:, 0], r[:, 1]).reshape(my, mx, cfg.c, cfg.c)
    rows = list(range(my // 2)) + list(range(n - my // 2, n))  # ky_r -> FFT row
    sel = xf[:, :, rows, :mx]                                  # (B, C, MY, MX)
    y = torch.einsum("bimk,mkio->bomk", sel, w)
    of = torch.zeros(v.shape[0], cfg.c, n, n // 2 + 1, dtype=xf.dtype)
    of[:, :, rows, :mx] = y
    return torch.fft.irfft2(of, s=(n, n))


def gelu(x):
    return 0.5 * x * (1.0 + torch.erf(x / np.sqrt(2.0)))


def forward(cfg, flat, omega):
    """omega (B, N, N) -> prediction (B, N, N); flat float64 parameters."""
    p = unflatten(cfg, flat)
    s = torch.sqrt(torch.mean(omega * omega, dim=(1, 2), keepdim=True))
    v = p["wl"][None, :, None, None] * (omega / s)[:, None] + p["bl"][None, :, None, None]
    for l in range(cfg.layers):
        k = spectral_conv(cfg, v, p["r"][l])
        mix = torch.einsum("oi,bixy->boxy", p["w"][l], v)
        pre = k + mix + p["b"][l][None, :, None, None]
        v = gelu(pre) if l < cfg.layers - 1 else pre
    h = gelu(torch.einsum("jc,bcxy->bjxy", p["w1"], v) + p["b1"][None, :, None, None])
    proj = torch.einsum("j,bjxy->bxy", p["w2"], h) + p["b2"]
    return omega + s * proj


def loss(pred, target):
    d = torch.sqrt(torch.sum((pred - target) ** 2, dim=(1, 2)))
    t = torch.sqrt(torch.sum(target ** 2, dim=(1, 2)))
    return torch.mean(d / t)


def run(cfg, params, omega, target=None):
    """numpy in, numpy out: prediction, and (loss, gradient) when target is given."""
    flat = torch.tensor(np.asarray(params, np.float64), requires_grad=target is not None)
    om = torch.tensor(np.asarray(omega, np.float64))
    pred = forward(cfg, flat, om)
    if target is None:
        return pred.detach().numpy()
    lv = loss(pred, torch.tensor(np.asarray(target, np.float64)))
    lv.backward()
    return pred.detach().numpy(), float(lv.detach()), flat.grad.numpy()
