"""GPU parity of the batched ensemble (tara_ens_*): every run of a mixed-size,
mixed-viscosity ensemble matches the same run done alone, graph replay matches
eager launches replay after replay, and AB2 runs reproduce the reference's
config-1 trajectory."""
import numpy as np
import pytest

import tara_amd as T
from conftest import rel_l2
from tara_amd.ic import random_phase_vorticity

pytestmark = pytest.mark.gpu


def _ensemble_inputs():
    n = [64, 128, 64, 256, 128, 64]
    nu = [1e-2, 1e-3, 5e-3, 1e-4, 2e-3, 1e-3]
    w0 = [random_phase_vorticity(m, seed=10 + r, k0=3.0 + r) for r, m in enumerate(n)]
    return n, nu, w0


def test_ensemble_matches_single_runs():
    print("ensemble vs single runs", flush=True)
    n, nu, w0 = _ensemble_inputs()
    dt = 5e-4
    e = T.Ensemble(n, nu, dt, T.SCHEME_RK4)
    for r, w in enumerate(w0):
        e.upload(r, w)
    e.advance(20)
    got = [e.download(r) for r in range(len(n))]
    e.close()
    for r, m in enumerate(n):
        s = T.CshSolver(m, m, nu[r], dt, T.SCHEME_RK4)
        s.upload(w0[r])
        s.advance(20)
        want = s.download()
        s.close()
        assert rel_l2(got[r], want) < 1e-14, r


def test_ensemble_with_512_bucket_matches_single_runs():
    """The benchmark ensemble's largest grid (512^2) batched with the others."""
    print("ensemble with a 512^2 bucket vs single runs", flush=True)
    n = [512, 128, 512, 256]
    nu = [1e-4, 1e-2, 1e-3, 3e-4]
    w0 = [random_phase_vorticity(m, seed=40 + r, k0=3.0 + 1.5 * r) for r, m in enumerate(n)]
    e = T.Ensemble(n, nu, 5e-4, T.SCHEME_RK4)
    for r, w in enumerate(w0):
        e.upload(r, w)
    e.advance(10)
    got = [e.download(r) for r in range(len(n))]
    e.close()
    for r, m in enumerate(n):
        s = T.CshSolver(m, m, nu[r], 5e-4, T.SCHEME_RK4)
        s.upload(w0[r])
        s.advance(10)
        want = s.download()
        s.close()
        assert rel_l2(got[r], want) < 1e-14, r


def test_graph_replays_match_eager():
    print("ensemble graph vs eager", flush=True)
    n, nu, w0 = _ensemble_inputs()
    eg = T.Ensemble(n, nu, 5e-4, T.SCHEME_RK4)
    ee = T.Ensemble(n, nu, 5e-4, T.SCHEME_RK4)
    ee.use_graph(False)
    for r, w in enumerate(w0):
        eg.upload(r, w)
        ee.upload(r, w)
    for rep in range(3):
        eg.advance(1)
        ee.advance(1)
        for r in range(len(n)):
            assert np.array_equal(eg.download(r), ee.download(r)), (rep, r)
    eg.close()
    ee.close()


@pytest.mark.parametrize("scheme", [T.SCHEME_RK4, T.SCHEME_AB2])
def test_fused_buckets_follow_uploads_and_replays(scheme):
    """Buckets of Ny >= 1024 run each forward y pass fused with the next inverse
    y pass, and every advance call starts with a stand-alone inverse y pass of
    the state.  A run whose state is uploaded between two advance calls must
    continue from the new state; every run must match the same sequence done
    alone, and graph replays must equal eager launches."""
    print("fused ensemble buckets", flush=True)
    n = [1024, 256, 1024, 2048]
    nu = [1e-3, 1e-3, 1e-4, 5e-4]
    dt = 2e-4
    w0 = [random_phase_vorticity(m, seed=40 + r, k0=4.0 + r) for r, m in enumerate(n)]
    w_new = random_phase_vorticity(1024, seed=99, k0=6.0)
    eg = T.Ensemble(n, nu, dt, scheme)
    ee = T.Ensemble(n, nu, dt, scheme)
    ee.use_graph(False)
    try:
        for e in (eg, ee):
            for r, w in enumerate(w0):
                e.upload(r, w)
            e.advance(3)
            e.upload(2, w_new)
            e.advance(2)
        got = [eg.download(r) for r in range(len(n))]
        for r in range(len(n)):
            assert np.array_equal(got[r], ee.download(r)), r
    finally:
        eg.close()
        ee.close()
    for r, m in enumerate(n):
        s = T.CshSolver(m, m, nu[r], dt, scheme)
        s.upload(w0[r])
        s.advance(3)
        if r == 2:
            s.upload(w_new)
        s.advance(2)
        want = s.download()
        s.close()
        err = rel_l2(got[r], want)
        print(f"  run {r} ({m}^2): ensemble vs single run rel L2 {err:.2e}", flush=True)
        assert err < 1e-14, r


@pytest.mark.parametrize("tag", ["asrun", "intent"])
def test_ensemble_ab2_config1_vs_reference(golden, tag):
    """Two copies of the config-1 run plus a differently sized run in one
    ensemble: the 64^2 runs reproduce the reference's 2000-step trajectory."""
    w0 = golden[f"c1_{tag}_w0"]
    e = T.Ensemble([64, 128, 64], [0.01, 0.01, 0.01], 1e-4, T.SCHEME_AB2)
    e.upload(0, w0)
    e.upload(1, random_phase_vorticity(128, seed=3))
    e.upload(2, w0)
    e.advance(2000)
    for r in (0, 2):
        assert rel_l2(e.download(r), golden[f"c1_{tag}_w2000"]) < 1e-10
    e.close()


def test_ensemble_tune_keeps_results():
    n, nu, w0 = _ensemble_inputs()
    a = T.Ensemble(n, nu, 5e-4, T.SCHEME_RK4)
    b = T.Ensemble(n, nu, 5e-4, T.SCHEME_RK4)
    b.tune(2)
    for r, w in enumerate(w0):
        a.upload(r, w)
        b.upload(r, w)
    a.advance(5)
    b.advance(5)
    for r in range(len(n)):
        assert rel_l2(b.download(r), a.download(r)) < 1e-14
    a.close()
    b.close()
