"""TEST INFRASTRUCTURE ONLY: CPU baseline timing for bench.py.

Times the reference arithmetic on the host cores on a bounded sample of a
benchmark workload and prints one JSON line.  Prefers the reference build
(_ref/libtara_ref.so: TARA's own routines + FFTW 3.3.9 compiled from the
reference sources); falls back to the C restatement (liboracle.so), labelled
"port".

  2d: classic RK4 over the reference derive2csh (the harness of
      ref_harness.f90) at N^2, nu = 1e-4, dt = 2e-4, on ONE core: the
      reference's csh path is serial by construction (no OpenMP in
      src/main2csh.f95, derive2csh.f95, ab2csh.f95).
  3d: the reference's AB2 loop (src/main3com.f95:536-617) over derive3com
      (diffusion nu = eta = 1e-3 through the one-constant patch) and ab3com at
      N^3, dt = 5e-4, with OpenMP and FFTW threads = the host's core count (the
      reference hard-wires 40, src/main3com.f95:194); one warm-up step (it
      touches the ~130 automatic arrays derive3com keeps on the stack) before
      the timed steps.

The reference keeps its automatic arrays on the stack, so this runs as a child
process with an unlimited stack (bench.py starts it that way).

    python -m oracle.cpu_baseline --case 2d --n 4096 --steps 3
    python -m oracle.cpu_baseline --case 3d --n 256 --steps 1
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import oracle as O  # noqa: E402


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def usable_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def field_2d(n):
    """Random-phase vorticity with E(k) ~ k^4 exp(-(k/4)^2) (the timing does not
    depend on the values; they only keep the transforms' inputs realistic)."""
    rng = np.random.default_rng(0)
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    w = np.zeros((n // 2 + 1, n), np.complex128, order="F")
    sel = (k > 0) & (k < 40)
    w[sel] = np.exp(2j * np.pi * rng.random(sel.sum())) * k[sel] ** 2 * np.exp(-(k[sel] / 4) ** 2)
    w[0, :] = 0
    return w


def field_3d(n):
    """rho = 1 and smooth velocity / magnetic fields (a few low modes), as
    half spectra in the reference layout."""
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij", sparse=True)
    ux = 0.5 * (np.sin(Z) + np.cos(Y)) + 0 * X
    uy = 0.5 * (np.sin(X) + np.cos(Z)) + 0 * Y
    uz = 0.5 * (np.sin(Y) + np.cos(X)) + 0 * Z
    bx, by, bz = 0.3 * np.cos(2 * Y) + 0 * X + 0 * Z, 0.3 * np.cos(2 * Z) + 0 * X + 0 * Y, 0.3 * np.cos(2 * X) + 0 * Y + 0 * Z
    rho = np.ones((n, n, n))
    energy = 1.0 + 0.5 * (ux ** 2 + uy ** 2 + uz ** 2) + 0.5 * (bx ** 2 + by ** 2 + bz ** 2)
    reals = [rho, ux, uy, uz, energy, bx, by, bz]
    return [np.asfortranarray(np.fft.rfftn(np.asarray(r, dtype=float).T).T) for r in reals]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", choices=("2d", "3d"), default="2d")
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=3)
    a = ap.parse_args()
    n = a.n
    kind = "reference" if O.ref_available() else "port"
    info = {"cpu_model": cpu_model(), "nproc": usable_cores()}
    if a.case == "2d":
        w = field_2d(n)
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_rk4csh_steps(w, n, 1e-4, 2e-4, a.steps)
        else:
            O.rk4csh_steps(w, n, 1e-4, 2e-4, a.steps)
        dt = time.perf_counter() - t0
        threads = 1
        pts = n * n
    else:
        threads = int(os.environ.get("OMP_NUM_THREADS", info["nproc"]))
        st = field_3d(n)
        L = 2.0 * np.pi
        if kind != "reference":
            raise SystemExit("the 3D CPU baseline needs the reference build")
        st, hist = O.ref_mhd_ab2_steps(st, n, L, L, L, 1e-3, 1e-3, 5e-4, 1, threads=threads)   # warm-up
        t0 = time.perf_counter()
        O.ref_mhd_ab2_steps(st, n, L, L, L, 1e-3, 1e-3, 5e-4, a.steps, threads=threads, hist=hist)
        dt = time.perf_counter() - t0
        pts = n ** 3
    print(json.dumps(dict(info, kind=kind, case=a.case, seconds=dt, n=n, steps=a.steps,
                          value=pts * a.steps / dt, cores=threads)))


if __name__ == "__main__":
    main()
