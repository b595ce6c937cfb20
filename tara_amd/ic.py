"""Seeded random-phase initial conditions of the benchmark workloads.

The reference's own initial conditions are deterministic profiles (shear layer
src/main2csh.f95:98-106; ABC flow src/main3com.f95:301-303); the benchmark
workloads instead start from random-phase spectra with a prescribed shell
spectrum E(k).  Built on the host once per run (not on the hot path).
"""
import numpy as np


def _half_plane_k(n):
    kx = np.arange(n // 2 + 1, dtype=float)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    return kx + 0 * ky, ky + 0 * kx


def random_phase_vorticity(n, seed=1, k0=4.0, energy=1.0):
    """Half spectrum (Nh, N) of omega with E(k) ~ k^4 exp(-(k/k0)^2), uniform
    phases, Hermitian on kx = 0, zero mean, Nyquist rows/columns zero, scaled to
    kinetic energy 1/2 <|u|^2> = energy (u = curl^-1 omega, Parseval with
    half-plane weights; unnormalised forward FFT convention)."""
    kx, ky = _half_plane_k(n)
    k = np.sqrt(kx * kx + ky * ky)
    rng = np.random.default_rng(seed)
    phase = np.exp(2j * np.pi * rng.random(k.shape))
    E = k**4 * np.exp(-(k / k0) ** 2)
    amp = np.zeros_like(k)
    nz = k > 0
    amp[nz] = np.sqrt(k[nz] * E[nz] / np.pi)
    w = amp * phase
    w[0, 0] = 0
    w[:, n // 2] = 0          # ky Nyquist row
    w[n // 2, :] = 0          # kx Nyquist column
    # Hermitian symmetry of the kx = 0 column: w(0, -ky) = conj w(0, ky)
    col = w[0, :].copy()
    for j in range(1, n // 2):
        col[n - j] = np.conj(col[j])
    w[0, :] = col
    wt = np.where(kx == 0, 1.0, 2.0)
    k2 = np.where(k > 0, k * k, 1.0)
    ke = 0.5 * np.sum(wt * np.abs(w) ** 2 / k2) / float(n) ** 4
    w *= np.sqrt(energy / ke)
    return np.asfortranarray(w)
