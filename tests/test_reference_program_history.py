"""The reference program's uninitialised AB2 history (src/main2csh.f95:64-65:
dt_omegak_old is allocated and first read at step 1, src/ab2csh.f95:12).
Run as built against the reference routines and FFTW (no GPU), once as is and
once with glibc's allocation perturbation (every allocation pre-filled with a
byte pattern): as is, its first snapshot equals the oracle's first AB2 step
with a zero history, the history every path here assumes; perturbed, the
first step differs, so the result does depend on the allocation's contents and
zero is what the program sees from fresh memory.  Needs the linked programs
(tests/fortran/Makefile.mains, built by __graft_entry__.build() where the
reference checkout is present)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

EXE = os.path.join(ROOT, "tests", "fortran", "_mains", "tara_snap_ref")
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="reference program not linked")


def _first_snapshot(tmp_path, name, env_extra):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_golden as G

    d = tmp_path / name
    (d / "output").mkdir(parents=True)
    (d / "input.ini").write_text(G.snap_ini(64, 4, 1))
    env = dict(os.environ, **env_extra)
    r = subprocess.run([EXE, "input.ini"], cwd=d, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    with open(d / "output" / "fort.00000100") as f:
        return np.array(f.read().split(), dtype=float).reshape(-1, 5)


def test_uninitialised_history_is_zero_in_fresh_memory(tmp_path):
    nx, nu, dt = 64, 0.01, 1e-3
    _, w = O.ic2csh(nx, nx, d_in=0.0234375)
    w1, _ = O.ab2csh_steps(w, np.zeros_like(w, order="F"), nx, nu, dt, 1)
    om = (O.c2r(w1, nx) / (nx * nx)).ravel()
    plain = _first_snapshot(tmp_path, "plain", {})
    assert np.max(np.abs(plain[:, 2] - om)) < 1e-12 * np.max(np.abs(om))
    perturbed = _first_snapshot(tmp_path, "perturbed", {"MALLOC_PERTURB_": "165"})
    diff = np.abs(perturbed[:, 2] - om)
    assert not np.all(np.isfinite(diff)) or np.max(diff) > 1e-3 * np.max(np.abs(om))
