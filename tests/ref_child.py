"""Runs one reference computation in a child process whose stack limit is
lifted before it starts: the reference's derive2csh keeps its 16 work arrays
as automatic (stack) arrays, 2.1 GB at 4096^2, and derive3com about 130 of
them, more than a running process's stack can grow to.  Test infrastructure
(the checker side only).

    python tests/ref_child.py derive2csh|rk4 in.npy out.npy nx nu [dt steps]
    python tests/ref_child.py mhd3 in.npy out.npy nx mu dt steps eta
"""
import os
import resource
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def run(kind, w, nx, nu, dt=0.0, steps=0, tmpdir="/tmp", timeout=600):
    """Computes the reference result for w in a child process; returns it."""
    import numpy as np

    src = os.path.join(tmpdir, f"tara_ref_in_{os.getpid()}.npy")
    dst = os.path.join(tmpdir, f"tara_ref_out_{os.getpid()}.npy")
    np.save(src, np.asfortranarray(w))

    def unlimited():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    try:
        subprocess.run([sys.executable, os.path.abspath(__file__), kind, src, dst, str(nx), repr(nu), repr(dt),
                        str(steps)], check=True, preexec_fn=unlimited, timeout=timeout)
        return np.asfortranarray(np.load(dst, allow_pickle=False))
    finally:
        for p in (src, dst):
            if os.path.exists(p):
                os.remove(p)


def run_mhd3(state, nx, mu, eta, dt, steps, tmpdir="/tmp", timeout=900):
    """The reference's 3D AB2 loop (derive3com with COMMON's mu and eta, then
    ab3com; zero history) over steps steps from the 8 fields of state, with
    OpenMP / FFTW threads = the cores this process may use, in a child
    process.  Returns (state, history) as lists of 8 fields; the history after
    one step is derive3com of the input."""
    import numpy as np

    src = os.path.join(tmpdir, f"tara_ref3_in_{os.getpid()}.npy")
    dst = os.path.join(tmpdir, f"tara_ref3_out_{os.getpid()}.npy")
    np.save(src, np.stack([np.asarray(f, dtype=np.complex128) for f in state], axis=-1))

    def unlimited():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    try:
        subprocess.run([sys.executable, os.path.abspath(__file__), "mhd3", src, dst, str(nx), repr(mu), repr(dt),
                        str(steps), repr(eta)], check=True, preexec_fn=unlimited, timeout=timeout)
        out = np.load(dst, allow_pickle=False)
        return ([np.asfortranarray(out[..., f]) for f in range(8)],
                [np.asfortranarray(out[..., 8 + f]) for f in range(8)])
    finally:
        for p in (src, dst):
            if os.path.exists(p):
                os.remove(p)


def main():
    import numpy as np

    sys.path.insert(0, os.path.dirname(HERE))
    import oracle as O

    kind, src, dst, nx, nu, dt, steps = sys.argv[1:8]
    nx, nu, dt, steps = int(nx), float(nu), float(dt), int(steps)
    if kind == "mhd3":
        eta = float(sys.argv[8])
        st = np.load(src, allow_pickle=False)
        state = [np.asfortranarray(st[..., f]) for f in range(8)]
        del st
        try:
            threads = len(os.sched_getaffinity(0))
        except AttributeError:
            threads = os.cpu_count() or 1
        lx = 2.0 * 3.14159265358979323846
        state, hist = O.ref_mhd_ab2_steps(state, nx, lx, lx, lx, nu, eta, dt, steps, threads=threads)
        np.save(dst, np.stack(state + hist, axis=-1))
        return
    w = np.asfortranarray(np.load(src, allow_pickle=False))
    if kind == "derive2csh":
        out = O.ref_derive2csh(w, nx, nu) if O.ref_available() else O.derive2csh(w, nx, nu)
    elif kind == "rk4":
        out = O.ref_rk4csh_steps(w, nx, nu, dt, steps) if O.ref_available() else O.rk4csh_steps(w, nx, nu, dt, steps)
    else:
        raise SystemExit(f"unknown computation {kind}")
    np.save(dst, out)


if __name__ == "__main__":
    main()
