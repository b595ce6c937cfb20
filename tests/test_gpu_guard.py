"""The non-finite guard of the resident steppers (no counterpart in the
reference, which never checks its state).  The condition is forced from the
host: a state with one NaN (or Inf) mode is uploaded, every kernel runs to its
normal end (NaN arithmetic faults nothing), and the sync reports it.

  * 2D, 2D compressible and 3D: sync and download return
    TARA_ERR_NONFINITE after the step; the flag clears when a new state is
    uploaded.
  * Ensemble: only the poisoned run is flagged, and the other runs of its
    bucket and of the other buckets match the same runs done alone.
"""
import numpy as np
import pytest

import tara_amd as T
from conftest import rel_l2
from tara_amd.ic import random_phase_mhd, random_phase_vorticity

pytestmark = pytest.mark.gpu


def _raises_nonfinite(fn):
    with pytest.raises(T.TaraError) as ei:
        fn()
    assert ei.value.code == T.ERR_NONFINITE, str(ei.value)


@pytest.mark.parametrize("scheme,n", [(T.SCHEME_RK4, 256), (T.SCHEME_AB2, 64), (T.SCHEME_RK4, 1024)])
def test_csh_guard(scheme, n):
    print(f"guard csh scheme {scheme} n {n}", flush=True)
    w = random_phase_vorticity(n, seed=3)
    s = T.CshSolver(n, n, 1e-3, 1e-4, scheme)
    s.upload(w)
    s.advance(3)
    s.sync()                      # a finite run passes
    bad = w.copy()
    bad[5, 7] = np.nan
    s.upload(bad)
    s.advance(1)
    _raises_nonfinite(s.sync)
    _raises_nonfinite(s.download)
    s.upload(w)                   # a new state clears the flag
    s.advance(2)
    s.sync()
    assert np.all(np.isfinite(s.download()))
    s.close()


def test_mhd_guard():
    print("guard mhd 32^3", flush=True)
    n = 32
    L = 2 * np.pi
    st = random_phase_mhd(n, seed=2)
    s = T.MhdSolver(n, n, n, T.MhdParams(L, L, L, 1e-3, 1e-3, 10.0, 5e-4))
    s.upload(st)
    s.advance(2)
    s.sync()
    bad = [a.copy() for a in st]
    bad[6][3, 2, 1] = np.inf
    s.upload(bad)
    s.advance(1)
    _raises_nonfinite(s.sync)
    s.ic_random(seed=1)
    s.advance(1)
    s.sync()
    s.close()


def test_ensemble_guard_isolates_the_failed_run():
    print("guard ensemble", flush=True)
    n = [128, 256, 128, 512]
    nu = [1e-3, 1e-4, 5e-3, 1e-4]
    w0 = [random_phase_vorticity(m, seed=60 + r) for r, m in enumerate(n)]
    e = T.Ensemble(n, nu, 5e-4, T.SCHEME_RK4)
    for r, w in enumerate(w0):
        e.upload(r, w)
    poisoned = w0[2].copy()
    poisoned[1, 1] = np.nan
    e.upload(2, poisoned)
    e.advance(5)
    _raises_nonfinite(e.sync)
    assert e.nonfinite() == [False, False, True, False]
    for r in (0, 1, 3):
        s = T.CshSolver(n[r], n[r], nu[r], 5e-4, T.SCHEME_RK4)
        s.upload(w0[r])
        s.advance(5)
        want = s.download()
        s.close()
        assert rel_l2(e.download(r), want) < 1e-14, r
    e.upload(2, w0[2])            # setting the run's state again clears its flag
    e.advance(1)
    e.sync()
    assert e.nonfinite() == [False] * 4
    e.close()


@pytest.mark.parametrize("family", ["coh", "com"])
def test_f2_guard(family):
    print(f"guard {family} 64^2", flush=True)
    n, L = 64, 2 * np.pi
    rng = np.random.default_rng(5)
    if family == "coh":
        prm = T.CohParams(np.pi, L, L, 2e-4, 2e-4, 3.77, 1e-4)
        reals = [1.0 + 0.02 * rng.standard_normal((n, n))] + [0.05 * rng.standard_normal((n, n)) for _ in range(2)]
    else:
        prm = T.ComParams(np.pi, L, L, 1e-3, 1e-4, 2.0, 0.0, 0.0, 1e-4)
        reals = [1.0 + 0.02 * rng.standard_normal((n, n))] + [0.05 * rng.standard_normal((n, n)) for _ in range(2)]
        reals += [2.0 + 0.02 * rng.standard_normal((n, n))] + [0.05 * rng.standard_normal((n, n)) for _ in range(2)]
    st = [T.fft_r2c(np.asfortranarray(a)) for a in reals]
    s = T.F2Solver(n, n, prm)
    s.upload(st)
    s.advance(3)
    s.sync()
    bad = [a.copy() for a in st]
    bad[1][4, 9] = np.nan
    s.upload(bad)
    s.advance(1)
    _raises_nonfinite(s.sync)
    _raises_nonfinite(s.download)
    s.upload(st)
    s.advance(2)
    s.sync()
    assert all(np.all(np.isfinite(a)) for a in s.download())
    s.close()
