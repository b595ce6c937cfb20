"""TEST INFRASTRUCTURE ONLY: CPU baseline timing for bench.py.

Times the reference arithmetic on the host cores on a bounded sample of the
benchmark workload and prints one JSON line.  Prefers the reference build
(_ref/libtara_ref.so: TARA's derive2csh + FFTW 3.3.9, serial as the reference's
csh path is, driven by the classic-RK4 harness of ref_harness.f90); falls back to
the C restatement (liboracle.so) labelled "port".

The reference's Fortran keeps its automatic arrays on the stack, so this runs
as a child process with an unlimited stack (bench.py starts it that way).

    python -m oracle.cpu_baseline --n 4096 --steps 1
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import oracle as O  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--nu", type=float, default=1e-4)
    ap.add_argument("--dt", type=float, default=2e-4)
    a = ap.parse_args()
    n = a.n
    rng = np.random.default_rng(0)
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    w = np.zeros((n // 2 + 1, n), np.complex128, order="F")
    sel = (k > 0) & (k < 40)
    w[sel] = np.exp(2j * np.pi * rng.random(sel.sum())) * k[sel] ** 2 * np.exp(-(k[sel] / 4) ** 2)
    w[0, :] = 0
    kind = "reference" if O.ref_available() else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        O.ref_rk4csh_steps(w, n, a.nu, a.dt, a.steps)
    else:
        O.rk4csh_steps(w, n, a.nu, a.dt, a.steps)
    dt = time.perf_counter() - t0
    print(json.dumps({"kind": kind, "seconds": dt, "n": n, "steps": a.steps,
                      "value": n * n * a.steps / dt, "cores": 1}))


if __name__ == "__main__":
    main()
