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


def random_phase_mhd(n, seed=1, k0=3.0, e_kin=1.0, e_mag=0.5, cs=10.0):
    """Eight (Nh, N, N) spectra (rho, rho ux, rho uy, rho uz, E, Bx, By, Bz) of a
    uniform-density state whose velocity and magnetic field are random-phase,
    solenoidal, with shell spectrum E(k) ~ k^4 exp(-2 (k/k0)^2), scaled to
    1/2 <|u|^2> = e_kin and 1/2 <|B|^2> = e_mag.  E = P + 1/2 rho u^2 + 1/2 B^2
    with P = 1 as the reference builds it (src/main3com.f95:324-325); it is
    frozen by the update and does not affect the evolution."""
    rng = np.random.default_rng(seed)
    kx = np.arange(n // 2 + 1, dtype=float)[:, None, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :, None]
    kz = np.fft.fftfreq(n, 1.0 / n)[None, None, :]
    K = [kx + 0 * ky + 0 * kz, 0 * kx + ky + 0 * kz, 0 * kx + 0 * ky + kz]
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    k = np.sqrt(k2)
    amp = np.where(k > 0, np.sqrt(k**4 * np.exp(-2.0 * (k / k0) ** 2)) / np.maximum(k, 1.0), 0.0)
    nyq = (np.abs(K[0]) == n // 2) | (np.abs(K[1]) == n // 2) | (np.abs(K[2]) == n // 2)
    amp = np.where(nyq, 0.0, amp)     # Nyquist planes zero: the projection stays exact

    def solenoidal(energy):
        comps = []
        for _ in range(3):
            spec = amp * np.exp(2j * np.pi * rng.random(k.shape))
            # back through real space so the half spectrum is Hermitian-consistent
            real = np.fft.irfftn(spec.T, s=(n, n, n), axes=(0, 1, 2)).T
            comps.append(np.fft.rfftn(real.T).T)
        kd = (K[0] * comps[0] + K[1] * comps[1] + K[2] * comps[2]) / np.where(k2 > 0, k2, 1.0)
        comps = [comps[c] - K[c] * kd for c in range(3)]
        reals = [np.fft.irfftn(c.T, s=(n, n, n), axes=(0, 1, 2)).T for c in comps]
        scale = np.sqrt(energy / (0.5 * np.mean(sum(r * r for r in reals))))
        return [r * scale for r in reals]

    u = solenoidal(e_kin)
    b = solenoidal(e_mag)
    rho = np.ones((n, n, n))
    energy = 1.0 + 0.5 * rho * sum(c * c for c in u) + 0.5 * sum(c * c for c in b)
    reals = [rho, rho * u[0], rho * u[1], rho * u[2], energy, b[0], b[1], b[2]]
    return [np.asfortranarray(np.fft.rfftn(r.T).T) for r in reals]
