"""Surrogate workflow around the FNO (tara_fno_*) and the FP64 ensemble
(tara_ens_*), all on the device:

* ``generate``  -- trajectories of the FP64 solver from seeded random-phase
  initial conditions, sampled every ``steps`` solver steps, as real-space
  float32 vorticity; consecutive samples are the training pairs.
* ``train``     -- Adam on the relative L2 error of one surrogate step, with a
  cosine learning-rate schedule, mini-batches drawn from the pairs.
* ``rollout``   -- surrogate rollouts re-checked by the FP64 solver every
  ``every`` steps (tara_fno_rollout), timed, with errors measured against FP64
  reference trajectories.

torch provides device memory only; every computation is a tara_amd kernel.
"""
import time

import numpy as np
import torch

from . import FNO, SCHEME_RK4, CshSolver, Ensemble


def generate(n, nu, dt, steps, seeds, nsamples, k0_range=(3.0, 8.0), energy=1.0, k0_seed=0):
    """FP64 trajectories of the derive2csh physics (RK4) for each seed.

    Returns (snaps, k0): snaps is device float32 [nsamples + 1][len(seeds)][n][n]
    with snaps[s][r] the real-space vorticity of seed r after s * steps steps;
    k0[r] is the peak wavenumber of its initial spectrum (uniform in k0_range,
    drawn from k0_seed)."""
    seeds = [int(v) for v in seeds]
    k0 = np.random.default_rng(k0_seed).uniform(k0_range[0], k0_range[1], len(seeds))
    ens = Ensemble([n] * len(seeds), [nu] * len(seeds), dt, SCHEME_RK4)
    try:
        for r, sd in enumerate(seeds):
            ens.ic_random(r, sd, float(k0[r]), energy)
        snaps = torch.empty((nsamples + 1, len(seeds), n, n), dtype=torch.float32, device="cuda")
        for s in range(nsamples + 1):
            for r in range(len(seeds)):
                ens.get_omega(r, snaps[s, r])
            if s < nsamples:
                ens.advance(steps)
        ens.sync()
    finally:
        ens.close()
    return snaps, k0


def train(net, snaps, epochs, batch, lr=2e-3, lr_min=5e-5, seed=0, log_every=0):
    """Trains on the pairs (snaps[s][r], snaps[s+1][r]); returns the mean loss of
    each epoch."""
    s1, nr, n, _ = snaps.shape
    npairs = (s1 - 1) * nr
    inp = snaps[:-1].reshape(npairs, n, n)
    tgt = snaps[1:].reshape(npairs, n, n)
    xb = torch.empty((batch, n, n), dtype=torch.float32, device="cuda")
    yb = torch.empty_like(xb)
    rng = np.random.default_rng(seed)
    per_epoch = npairs // batch
    total = epochs * per_epoch
    hist, it = [], 0
    for e in range(epochs):
        perm = rng.permutation(npairs)
        acc = 0.0
        for i in range(per_epoch):
            net.gather_pairs(perm[i * batch:(i + 1) * batch], inp, tgt, xb, yb)
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * it / max(total - 1, 1)))
            loss = net.train_step(batch, xb, yb, lr_t)
            acc += loss
            it += 1
            if log_every and it % log_every == 0:
                print(f"  step {it}/{total} loss {loss:.4f} lr {lr_t:.2e}", flush=True)
        hist.append(acc / max(per_epoch, 1))
    return hist


def one_step_error(net, snaps, batch=16):
    """Mean relative L2 error of a single surrogate step over the pairs of snaps,
    and that of persistence (omega(t + T) = omega(t)) on the same pairs."""
    s1, nr, n, _ = snaps.shape
    npairs = (s1 - 1) * nr
    inp = snaps[:-1].reshape(npairs, n, n)
    tgt = snaps[1:].reshape(npairs, n, n)
    out = torch.empty((batch, n, n), dtype=torch.float32, device="cuda")
    e_net, e_id = [], []
    for i in range(0, npairs - batch + 1, batch):
        net.forward(batch, inp[i:i + batch], out)
        e_net.append(net.rel_err(batch, out, tgt[i:i + batch]))
        e_id.append(net.rel_err(batch, inp[i:i + batch], tgt[i:i + batch]))
    return float(np.mean(np.concatenate(e_net))), float(np.mean(np.concatenate(e_id)))


def rollout(net, ref, nu, dt, steps, every=10, tol=0.05, graph=True, reps=1, batched=True):
    """Checked surrogate rollouts of each trajectory in ref (device float32
    [nsteps + 1][R][n][n], FP64 reference states every `steps` solver steps),
    one trajectory at a time.  batched: the checks run as one ensemble after
    the rollout (else one ensemble per check, started as inputs appear).

    Returns a dict with, per trajectory: the relative L2 error of the final
    state against the FP64 reference, the largest error over all steps, the
    checker's errors, the number of solver take-overs, and the wall time of
    the rollout (best of `reps`)."""
    s1, nr, n, _ = ref.shape
    nsteps = s1 - 1
    nchk = nsteps // every
    if nchk > 1 and batched:
        checkers = [Ensemble([n] * nchk, [nu] * nchk, dt, SCHEME_RK4)]
    else:
        checkers = [Ensemble([n], [nu], dt, SCHEME_RK4) for _ in range(nchk)]
    anchor = Ensemble([n], [nu], dt, SCHEME_RK4) if nchk > 0 else None
    traj = torch.empty((s1, 1, n, n), dtype=torch.float32, device="cuda")
    out = {"final": [], "max": [], "check_err": [], "fallbacks": [], "ms": []}
    try:
        for r in range(nr):
            best = None
            for _ in range(max(reps, 1)):
                traj[0, 0].copy_(ref[0, r])
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                err, nf = net.rollout(1, traj, nsteps, every, checkers, steps, tol, graph, anchor=anchor)
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) * 1e3
                best = ms if best is None else min(best, ms)
            errs = net.rel_err(1, traj[nsteps], ref[nsteps, r:r + 1])
            step_err = [float(net.rel_err(1, traj[t], ref[t, r:r + 1])[0]) for t in range(1, s1)]
            out["final"].append(float(errs[0]))
            out["max"].append(max(step_err))
            out["check_err"].append(err[:, 0].tolist())
            out["fallbacks"].append(nf)
            out["ms"].append(best)
    finally:
        for c in checkers + ([anchor] if anchor is not None else []):
            c.close()
    return out


def solver_time(n, nu, dt, total_steps, seed=1, k0=5.0):
    """Wall time (ms) of the plain FP64 GPU solver advancing one run by
    total_steps RK4 steps (the single-run device-resident path)."""
    s = CshSolver(n, n, nu=nu, dt=dt, scheme=SCHEME_RK4)
    try:
        s.ic_random(seed=seed, k0=k0)
        s.advance(20)
        s.sync()
        t0 = time.perf_counter()
        s.advance(total_steps)
        s.sync()
        return (time.perf_counter() - t0) * 1e3
    finally:
        s.close()


def _main(argv=None):
    """Times network steps on random fields (for profiling):
    python -m tara_amd.surrogate [--batch B] [--reps R] [--train] [--n N]"""
    import argparse

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--train", action="store_true", help="time training steps instead of inference")
    a = ap.parse_args(argv)
    net = FNO(a.n, 64, 16, 32, 4, max_batch=a.batch, seed=1)
    om = torch.randn((a.batch, a.n, a.n), device="cuda")
    tgt = 0.9 * om
    out = torch.empty_like(om)
    step = (lambda: net.train_step(a.batch, om, tgt, 1e-4)) if a.train else (lambda: net.forward(a.batch, om, out))
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.reps
    print(f"{'train' if a.train else 'forward'} batch {a.batch}: {ms:.3f} ms per step", flush=True)


if __name__ == "__main__":
    _main()
