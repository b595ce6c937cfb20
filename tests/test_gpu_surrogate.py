"""The surrogate workflow on a small grid: FP64 ensemble data, training on the
device, and rollouts re-checked by the FP64 solver.

* The trained network's one-step error on held-out seeds is below that of the
  method without it (persistence, omega(t + T) = omega(t)) and of the untrained
  network, by at least a fifth.
* The checker's error for a step equals the relative L2 distance between the
  surrogate's step and an independent FP64 solve from the same state.
* With a tolerance no step meets, every check hands the trajectory to the
  FP64 solver: each checked state is the plain solver's state at that step,
  advanced from the initial field (the anchor), not a correction of a drifted
  surrogate state."""
import numpy as np
import pytest

import tara_amd as T
from tara_amd import surrogate as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N, NU, DT, STEPS = 64, 1e-3, 5e-4, 100


@pytest.fixture(scope="module")
def data():
    print("surrogate: generating 64x64 trajectories", flush=True)
    train, _ = S.generate(N, NU, DT, STEPS, range(1, 65), 20, k0_range=(3.0, 5.0))
    held, _ = S.generate(N, NU, DT, STEPS, range(9001, 9005), 20, k0_range=(3.0, 5.0), k0_seed=7)
    return train, held


@pytest.fixture(scope="module")
def trained(data):
    train, held = data
    net = T.FNO(N, 32, 12, 24, 4, max_batch=16, seed=1)
    e_untrained = S.one_step_error(net, held)[0]
    hist = S.train(net, train, 20, 16, lr=1e-3, lr_min=1e-4)
    print(f"  losses {np.round(hist, 4)}", flush=True)
    return net, e_untrained


def test_training_beats_persistence(data, trained):
    _, held = data
    net, e_untrained = trained
    e_net, e_id = S.one_step_error(net, held)
    print(f"  one-step error: trained {e_net:.4f}, untrained {e_untrained:.4f}, persistence {e_id:.4f}", flush=True)
    # measured: trained 0.176, untrained and persistence 0.271
    assert e_net < 0.8 * e_id
    assert e_net < 0.8 * e_untrained


def _solve(state, nsteps):
    """FP64 solver advance of one real-space float32 field (independent of the rollout's checkers)."""
    e = T.Ensemble([N], [NU], DT, T.SCHEME_RK4)
    out = torch.empty_like(state)
    e.set_omega(0, state)
    e.advance(nsteps)
    e.get_omega(0, out)
    e.sync()
    e.close()
    return out


def test_checker_errors_match_independent_solves(data, trained):
    _, held = data
    net, _ = trained
    nsteps, every = 20, 5
    checkers = [T.Ensemble([N], [NU], DT, T.SCHEME_RK4) for _ in range(nsteps // every)]
    anchor = T.Ensemble([N], [NU], DT, T.SCHEME_RK4)
    traj = torch.empty((nsteps + 1, 1, N, N), device="cuda")
    traj[0, 0].copy_(held[0, 0])
    err, nf = net.rollout(1, traj, nsteps, every, checkers, STEPS, tol=1e9, anchor=anchor)
    assert nf == 0
    # without take-overs the rollout is the plain network iteration
    plain = torch.empty_like(traj)
    plain[0].copy_(traj[0])
    net.rollout(1, plain, nsteps, nsteps + 1, (), STEPS, 1e9)
    assert torch.equal(plain, traj)
    for j in range(nsteps // every):
        t = (j + 1) * every
        ref = _solve(traj[t - 1, 0], STEPS)
        want = float(torch.linalg.norm(traj[t, 0] - ref) / torch.linalg.norm(ref))
        assert abs(err[j, 0] - want) < 1e-4 * max(want, 1e-3), (j, err[j, 0], want)
    for c in checkers + [anchor]:
        c.close()


@pytest.mark.parametrize("batched", [False, True])
def test_failed_checks_hand_the_trajectory_to_the_solver(data, trained, batched):
    """tol = 0: every check fails, and every checked step of the rollout is
    the plain FP64 solver's state at that step from traj[0]: the take-overs
    follow one FP64 trajectory instead of correcting drifted surrogate states."""
    _, held = data
    net, _ = trained
    nsteps, every = 20, 5
    nchk = nsteps // every
    if batched:
        checkers = [T.Ensemble([N] * nchk, [NU] * nchk, DT, T.SCHEME_RK4)]
    else:
        checkers = [T.Ensemble([N], [NU], DT, T.SCHEME_RK4) for _ in range(nchk)]
    anchor = T.Ensemble([N], [NU], DT, T.SCHEME_RK4)
    traj = torch.empty((nsteps + 1, 1, N, N), device="cuda")
    traj[0, 0].copy_(held[0, 0])
    err, nf = net.rollout(1, traj, nsteps, every, checkers, STEPS, tol=0.0, anchor=anchor)
    assert nf == nchk
    assert np.all(err > 0)
    plain = T.Ensemble([N], [NU], DT, T.SCHEME_RK4)
    plain.set_omega(0, traj[0, 0])
    out = torch.empty_like(traj[0, 0])
    for j in range(nchk):
        plain.advance(every * STEPS)
        plain.get_omega(0, out)
        plain.sync()
        t = (j + 1) * every
        assert torch.equal(traj[t, 0], out), j
        # and the surrogate runs on from there
        assert not torch.equal(traj[t - 1, 0], traj[t, 0])
    for c in checkers + [anchor, plain]:
        c.close()


def test_ensemble_graph_captured_by_advance_zero():
    """advance(0) captures the step graph without stepping (a caller can
    capture before other work is in flight) and leaves the state unchanged."""
    e = T.Ensemble([N, N], [NU, 2 * NU], DT, T.SCHEME_RK4)
    e.ic_random(0, 3)
    e.ic_random(1, 4)
    before = [e.download(r) for r in range(2)]
    assert e.graph_nodes() == (0, 0)
    e.advance(0)
    nodes, kernels = e.graph_nodes()
    assert kernels > 0 and nodes >= kernels
    assert all(np.array_equal(e.download(r), before[r]) for r in range(2))
    e.close()


def test_batched_checker_matches_separate_checkers(data, trained):
    """One ensemble holding every check gives the same errors, take-overs and
    trajectory as one ensemble per check."""
    _, held = data
    net, _ = trained
    nsteps, every = 20, 5
    outs = []
    for batched in (False, True):
        if batched:
            checkers = [T.Ensemble([N] * (nsteps // every), [NU] * (nsteps // every), DT, T.SCHEME_RK4)]
        else:
            checkers = [T.Ensemble([N], [NU], DT, T.SCHEME_RK4) for _ in range(nsteps // every)]
        anchor = T.Ensemble([N], [NU], DT, T.SCHEME_RK4)
        traj = torch.empty((nsteps + 1, 1, N, N), device="cuda")
        traj[0, 0].copy_(held[0, 1])
        err, nf = net.rollout(1, traj, nsteps, every, checkers, STEPS, tol=0.0, anchor=anchor)
        outs.append((err, nf, traj.cpu().numpy()))
        for c in checkers + [anchor]:
            c.close()
    (e0, n0, t0), (e1, n1, t1) = outs
    assert n0 == n1 == nsteps // every
    assert np.allclose(e0, e1, rtol=1e-6, atol=0)
    assert np.allclose(t0, t1, rtol=0, atol=1e-6 * np.abs(t0).max())
