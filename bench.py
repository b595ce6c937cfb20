"""Benchmark: FP64 grid-point steps per second of the 2D incompressible
pseudo-spectral solver, RK4, 4096^2 (nu = 1e-4, dt = 2e-4, random-phase
vorticity with E(k) ~ k^4 exp(-(k/4)^2), unit kinetic energy).

One "step" = one full RK4 step (4 right-hand sides, each: inverse y pass with
the spectral multiplies, x pass with c2r / products / r2c, forward y pass with
dealias + viscous term + the fused stage update) on one GPU, state resident in
HBM.  value = Nx*Ny*steps / seconds.

Prints one JSON line.  Also reports the roofline of the dominant kernel
(algorithmic bytes / average HIP-event duration, against 8 TB/s HBM) and CPU
baselines timed in child processes: the reference's own derive2csh + FFTW on
one host core (it is serial by construction), three RK4 steps at the same
size; and for the 3D line the reference's derive3com + ab3com with OpenMP and
FFTW threads = the host's cores, three AB2 steps at 256^3.
"""
import argparse
import json
import math
import os
import resource
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X vector FP64

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and, in a
# separate run, WRITE_SIZE over this script (--steps 2 --warmup 1
# --no-cpu-baseline --no-ensemble --no-fno), each scaled by what the counter reads for a device copy
# of known size (tara_amd/calib/copy_probe: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).
MEASURED_TRAFFIC = {
    ("k_csh_yinv", 4096): 7.1266e8,
    ("k_csh_yfwd", 4096): 6.7424e8,      # mean over the four RK4 stage variants
    ("k_csh_yfwd_inv", 4096): 1.2985e9,  # mean over the four RK4 stage variants
    # k_csh_x per x-pass shape (the autotuner's choice, see tara_csh_tune)
    ("k_csh_x", 4096, 0): 9.4956e8,
    ("k_csh_x", 4096, 1): 7.1194e8,
    ("k_csh_x", 4096, 2): 6.7919e8,
    ("k_csh_x", 4096, 3): 2.7163e9,
    ("k_csh_x", 4096, 4): 7.1465e8,
    ("k_csh_x", 4096, 5): 6.9267e8,
    ("k_mhd_col_inverse_zy", 256): 3.9932e9,   # two launches (z, then y), 1.9966e9 each
    ("k_mhd_x", 256): 2.9069e9,
    ("k_mhd_col_forward", 256): 3.9931e9,
    ("k_mhd_zfwd", 256): 6.1173e9,
}
METRIC = "grid-pt steps/s FP64 (2D 4096², 3D 256³ MHD); FNO sim-time/s at ≤5% L2 err"


def model_2d(n, nsteps):
    """Algorithmic bytes of the 2D RK4 path at n^2 for one advance call of
    nsteps steps: per timing class (mean per launch), class names, and bytes
    per step.  S = one half spectrum.  The inverse y pass reads 1 field and
    writes 4; the x pass reads 4 and writes 1; the forward y pass of RK4 stages
    1..4 moves 4, 6, 6 and 4 fields.  From Ny = 1024 each forward y pass is fused
    with the next stage's inverse y pass (timing class 2, + 4 S written, no
    read), except the call's last, and class 0 is the call's one stand-alone
    inverse y pass: (36 nsteps - 4) S in class 2, 56 S per step.  Unfused:
    5 S per launch in every class, 60 S per step."""
    S = 16.0 * (n // 2 + 1) * n
    if n >= 1024:
        return ({0: 5 * S, 1: 5 * S, 2: (36 * nsteps - 4) * S / (4 * nsteps)},
                {0: "k_csh_yinv", 1: "k_csh_x", 2: "k_csh_yfwd_inv"}, 56 * S)
    return {0: 5 * S, 1: 5 * S, 2: 5 * S}, {0: "k_csh_yinv", 1: "k_csh_x", 2: "k_csh_yfwd"}, 60 * S


def cpu_baseline(case, n, steps):
    """The reference on this host's cores in a child process (unlimited stack:
    the reference keeps its work arrays there): 2D on one core (its csh path
    is serial), 3D with OpenMP / FFTW threads = the cores this process may use."""
    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    env = dict(os.environ, OMP_NUM_THREADS="1" if case == "2d" else str(cores))
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--case", case, "--n", str(n), "--steps",
                        str(steps)], cwd=ROOT, capture_output=True, text=True, preexec_fn=unlimited_stack, env=env,
                       timeout=600)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-2000:])
        return None
    d = json.loads(r.stdout.strip().splitlines()[-1])
    src = ("TARA Fortran + FFTW 3.3.9 built from the reference sources" if d["kind"] == "reference"
           else "C restatement (not FFTW)")
    if case == "2d":
        sample = (f"{steps} RK4 steps (4 derive2csh calls + stage updates each) at {n}^2, "
                  f"{d['seconds']:.1f} s on 1 host core, {src}")
    else:
        sample = (f"{steps} AB2 step(s) (derive3com + ab3com, nu = eta = 1e-3 via the one-constant patch) at "
                  f"{n}^3 after one warm-up step, {d['seconds']:.1f} s on {d['cores']} threads, {src}")
    return {"value": d["value"], "unit": "grid-pt steps/s", "cores": d["cores"], "kind": d["kind"],
            "sample": sample, "cpu_model": d["cpu_model"], "nproc": d["nproc"], "threads": d["cores"]}


def device_copy_gbps(mib=1024):
    """HBM copy bandwidth (read + write) of a plain device copy of mib MiB per
    buffer, larger than the last-level cache, from tara_amd/calib/copy_probe
    in a child process (the better of its two launch shapes): a practical
    ceiling beside the 8 TB/s peak."""
    exe = os.path.join(ROOT, "tara_amd", "calib", "copy_probe")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe, str(mib)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-2000:])
        return None
    try:
        return float(r.stdout.split("best ")[1].split(" GB/s")[0])
    except (ValueError, IndexError):
        return None


def bench_2d(n, steps, warmup, nu=1e-4, dt=2e-4):
    """Secondary line: the same RK4 step at another grid (the configuration's
    other size, 1024^2): value, per-kernel times and the model rate."""
    import tara_amd as T
    from tara_amd.ic import random_phase_vorticity

    s = T.CshSolver(n, n, nu, dt, T.SCHEME_RK4)
    s.upload(random_phase_vorticity(n, seed=1, k0=4.0))
    xshape, _ = s.tune(reps=5)
    s.advance(warmup)
    s.sync()
    reps = steps * max(1, (4096 // n) ** 2 // 4)
    t0 = time.perf_counter()
    s.advance(reps)
    s.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    s.set_timing(True)
    s.advance(steps)
    s.sync()
    counts, kms = s.kernel_stats(3)
    s.close()
    model, names, step_bytes = model_2d(n, steps)
    kernels = {names[c]: {"avg_ms": round(kms[c] / counts[c], 4), "launches": counts[c],
                          "GBps": round(model[c] / (kms[c] / counts[c] * 1e-3) / 1e9, 1)}
               for c in range(3) if counts[c]}
    return {"workload": f"2D incompressible hydro, RK4, {n}^2, FP64", "value": n * n / (ms * 1e-3),
            "unit": "grid-pt steps/s", "ms_per_step": ms, "steps": reps, "xpass_shape": xshape, "kernels": kernels,
            "step_model_bytes": step_bytes, "step_model_GBps": round(step_bytes / (ms * 1e-3) / 1e9, 1)}


def bench_mhd(steps, warmup, n=256):
    """Secondary line: 3D compressible MHD, AB2, 256^3, nu = eta = 1e-3,
    dt = 5e-4 (derive3com + ab3com on the GPU, state resident)."""
    import tara_amd as T
    from tara_amd.ic import random_phase_mhd

    L = 2.0 * 3.14159265358979323846
    prm = T.MhdParams(L, L, L, 1e-3, 1e-3, 10.0, 5e-4)
    s = T.MhdSolver(n, n, n, prm)
    s.upload(random_phase_mhd(n, seed=1))
    s.advance(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.advance(steps)
    s.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    s.set_timing(True)
    s.advance(steps)
    s.sync()
    counts, kms = s.kernel_stats(4)
    s.close()
    S3 = 16.0 * (n // 2 + 1) * n * n
    # algorithmic bytes per timed region: inverse z and y passes (two launches)
    # 7 fields in + 7 out each;
    # x pass 7 in, 14 out; forward y 14 + 14; forward z reads 14 fluxes + 7 state
    # + 7 history and writes 7 state + 7 history.  119 S3 per step.
    names = ["k_mhd_col_inverse_zy", "k_mhd_x", "k_mhd_col_forward", "k_mhd_zfwd"]
    model = [28 * S3, 21 * S3, 28 * S3, 42 * S3]
    avg = {c: kms[c] / counts[c] for c in range(4) if counts[c]}
    dom = max(avg, key=lambda c: kms[c])
    achieved = model[dom] / (avg[dom] * 1e-3) / 1e9
    return {"workload": f"3D compressible isothermal MHD, AB2, {n}^3, FP64, nu = eta = 1e-3, dt = 5e-4",
            "value": n ** 3 / (ms * 1e-3), "unit": "grid-pt steps/s", "ms_per_step": ms,
            "step_model_bytes": 119 * S3, "step_model_GBps": round(119 * S3 / (ms * 1e-3) / 1e9, 1),
            "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": MEASURED_TRAFFIC.get((names[dom], n))},
            "kernels": {names[c]: {"avg_ms": round(avg[c], 4), "launches": counts[c],
                                   "GBps": round(model[c] / (avg[c] * 1e-3) / 1e9, 1)} for c in avg}}


def bench_ensemble(steps, warmup, nruns=512, seed=7):
    """Secondary line: batched ensemble of 512 independent 2D runs, grids drawn
    uniformly from {128, 256, 512}^2, nu log-uniform in [1e-4, 1e-2], k0 uniform
    in [3, 8], dt = 5e-4, RK4, one hipGraph replay per step; against the same
    512 runs done one after another on the single-run solver."""
    import numpy as np

    import tara_amd as T
    from tara_amd.ic import random_phase_vorticity

    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in rng.choice([128, 256, 512], size=nruns)]
    nus = 10.0 ** rng.uniform(-4, -2, size=nruns)
    k0s = rng.uniform(3, 8, size=nruns)
    e = T.Ensemble(sizes, nus, 5e-4, T.SCHEME_RK4)
    for r in range(nruns):
        e.upload(r, random_phase_vorticity(sizes[r], seed=1000 + r, k0=k0s[r]))
    e.tune(3)
    e.advance(warmup)
    e.sync()
    t0 = time.perf_counter()
    e.advance(steps)
    e.sync()
    t_ens = (time.perf_counter() - t0) / steps
    e.close()
    seq = 0.0
    singles = {}
    for n in (128, 256, 512):
        s = T.CshSolver(n, n, 1e-3, 5e-4, T.SCHEME_RK4)
        s.upload(random_phase_vorticity(n, seed=5))
        s.tune(reps=3)
        s.advance(warmup)
        s.sync()
        reps = 4 * steps
        t0 = time.perf_counter()
        s.advance(reps)
        s.sync()
        singles[n] = (time.perf_counter() - t0) / reps
        s.close()
        seq += sizes.count(n) * singles[n]
    pts = float(sum(n * n for n in sizes))
    S_all = sum(16.0 * (n // 2 + 1) * n for n in sizes)
    return {"workload": f"ensemble of {nruns} independent 2D runs, grids U{{128,256,512}}^2, nu log-U[1e-4,1e-2], "
                        "RK4, dt 5e-4, FP64, hipGraph replay",
            "value": pts / t_ens, "unit": "grid-pt steps/s", "ms_per_step": t_ens * 1e3,
            "runs_per_size": {n: sizes.count(n) for n in (128, 256, 512)},
            "sequential_value": pts / seq, "sequential_ms_per_step": seq * 1e3,
            "single_run_ms_per_step": {n: round(v * 1e3, 4) for n, v in singles.items()},
            "speedup_vs_sequential": seq / t_ens,
            "step_model_bytes": 60 * S_all, "step_model_GBps": round(60 * S_all / t_ens / 1e9, 1),
            # FP64 transform work: 4 RHS x 5 real 2D transforms of 2.5 N log2 N flops per run
            "fp64_fft_tflops": round(sum(20 * 2.5 * n * n * math.log2(n * n) for n in sizes) / t_ens / 1e12, 2)}


def fno_forward_flops(n, c, mx, my, layers):
    """Algorithmic FLOPs of one network step for one sample (the products of
    tara_amd/csrc/fno.hip's forward: truncated DFTs, per-mode contraction,
    inverse DFTs, channel mixing, projection)."""
    np_ = n * n
    layer = (2 * 2 * mx * n * np_ // n * c        # x transform: (2MX x n) (n x n) per channel
             + 2 * 2 * my * 2 * n * mx * c        # y transform
             + 2 * (2 * c) * (2 * c) * mx * my    # per-mode contraction
             + 2 * 2 * n * 2 * my * mx * c        # inverse y
             + 2 * c * n * 2 * mx * n             # inverse x
             + 2 * c * c * np_)                   # channel mixing
    return layers * layer + 2 * 2 * c * c * np_ + 2 * 2 * c * np_


def bench_fno(epochs, nseeds=400, nheld=8):
    """Secondary line: the FNO surrogate of 2D hydro at 256^2, nu 1e-3, one
    network step = 100 solver steps (dt 5e-4).  Trains on nseeds x 50 pairs
    from the FP64 ensemble (k0 uniform in [3, 8]), then rolls out 50 steps from
    nheld held-out seeds, re-checked by the FP64 solver every 10th step with
    take-over above 5 % relative L2 error; errors are measured against FP64
    reference trajectories, speed against the single-run FP64 solver."""
    import numpy as np
    import torch

    import tara_amd as T
    from tara_amd import surrogate as S

    n, nu, dt, steps, width, mx, my, layers, batch = 256, 1e-3, 5e-4, 100, 64, 16, 32, 4, 16
    t0 = time.perf_counter()
    snaps, _ = S.generate(n, nu, dt, steps, range(1, nseeds + 1), 50)
    held, _ = S.generate(n, nu, dt, steps, range(100001, 100001 + nheld), 50, k0_seed=99)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    net = T.FNO(n, width, mx, my, layers, max_batch=batch, seed=3)
    xb = snaps[0, :batch].contiguous()
    net.train_step(batch, xb, snaps[1, :batch].contiguous(), 0.0)  # warm
    t0 = time.perf_counter()
    hist = S.train(net, snaps, epochs, batch, lr=1e-3, lr_min=5e-5)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    npairs = 50 * nseeds
    e_net, e_id = S.one_step_error(net, held)
    out = torch.empty((batch, n, n), device="cuda")
    net.forward(batch, snaps[0, :batch], out)
    reps = 20
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(batch, snaps[0, :batch], out)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / reps
    chk = S.rollout(net, held, nu, dt, steps, every=10, tol=0.05, reps=2)
    free = S.rollout(net, held, nu, dt, steps, every=51, tol=0.05, reps=2)
    # speed only: every check runs, none takes over (tolerance above any error);
    # checks batched into one ensemble, and one ensemble per check for comparison
    cpass = S.rollout(net, held, nu, dt, steps, every=10, tol=1e9, reps=2)
    cpass_sep = S.rollout(net, held, nu, dt, steps, every=10, tol=1e9, reps=2, batched=False)
    nsteps = held.shape[0] - 1
    sim_time = nsteps * steps * dt
    solver_ms = S.solver_time(n, nu, dt, nsteps * steps)
    med = lambda v: float(np.median(v))
    flops = fno_forward_flops(n, width, mx, my, layers)
    return {
        "workload": f"FNO surrogate, 2D hydro {n}^2, nu {nu}, step = {steps} solver steps (dt {dt}); "
                    f"{layers} Fourier layers, width {width}, modes kx<{mx} x |ky|<={my // 2}, bf16 MFMA",
        "unit": "sim-time/s",
        "solver_value": sim_time / (solver_ms * 1e-3),
        "checked_value": sim_time / (med(chk["ms"]) * 1e-3),
        "unchecked_value": sim_time / (med(free["ms"]) * 1e-3),
        "speedup_checked": solver_ms / med(chk["ms"]),
        "speedup_unchecked": solver_ms / med(free["ms"]),
        "rollout_steps": nsteps, "check_every": 10, "tol": 0.05,
        "checked_final_rel_l2": [round(v, 4) for v in chk["final"]],
        "checked_fallbacks": chk["fallbacks"],
        "unchecked_final_rel_l2": [round(v, 4) for v in free["final"]],
        "one_step_rel_l2": e_net, "persistence_one_step_rel_l2": e_id,
        "target_met": bool(med(chk["final"]) <= 0.05 and solver_ms / med(chk["ms"]) >= 20),
        "train_pairs": npairs, "epochs": epochs, "epoch_loss": [round(v, 4) for v in hist],
        "train_s": t_train, "train_pairs_per_s": epochs * (npairs // batch) * batch / t_train,
        "datagen_s": t_gen,
        "forward_ms_batch16": fwd_ms,
        "forward_tflops_batch16": flops * batch / (fwd_ms * 1e-3) / 1e12,
        "forward_gflop_per_sample": flops / 1e9,
        "solver_ms_per_rollout": solver_ms,
        "checked_ms_per_rollout": med(chk["ms"]), "unchecked_ms_per_rollout": med(free["ms"]),
        "checks_without_takeover_ms_per_rollout": med(cpass["ms"]),
        "speedup_checks_without_takeover": solver_ms / med(cpass["ms"]),
        "checks_without_takeover_separate_ms_per_rollout": med(cpass_sep["ms"]),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-3d", action="store_true", help="skip the secondary 3D MHD measurement")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the secondary ensemble measurement")
    ap.add_argument("--no-fno", action="store_true", help="skip the secondary FNO surrogate measurement")
    ap.add_argument("--fno-epochs", type=int, default=3, help="training epochs of the FNO leg")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)
    n = a.n
    nu, dt = 1e-4, 2e-4

    # CPU baselines first, in child processes, before this process touches the GPU.
    cpu = None if a.no_cpu_baseline else cpu_baseline("2d", n, 3)
    cpu3 = None if (a.no_cpu_baseline or a.no_3d) else cpu_baseline("3d", 256, 3)
    copy_gbps = device_copy_gbps()
    if not a.no_fno:
        # torch (device memory of the surrogate leg) is loaded before the HIP
        # runtime starts in this process: loaded after, it finds no device.
        import torch  # noqa: F401

    import tara_amd as T
    from tara_amd.ic import random_phase_vorticity

    w0 = random_phase_vorticity(n, seed=1, k0=4.0)
    s = T.CshSolver(n, n, nu, dt, T.SCHEME_RK4)
    s.upload(w0)
    # the autotuner picks the x-pass shape on this grid before anything is timed
    xshape, _ = s.tune(reps=5)
    s.advance(a.warmup)
    s.sync()
    t0 = time.perf_counter()
    s.advance(a.steps)
    s.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps

    # Same region again with HIP events around every launch (per kernel class).
    s.set_timing(True)
    s.advance(a.steps)
    s.sync()
    counts, ms = s.kernel_stats(3)
    s.set_timing(False)
    w = s.download()
    s.close()
    finite = bool(abs(w).max() < 1e300)

    model, names, step_bytes = model_2d(n, a.steps)  # bytes per launch and per step (see model_2d)
    avg = {c: ms[c] / counts[c] for c in range(3) if counts[c]}
    dom = max(avg, key=lambda c: ms[c])
    achieved = model[dom] / (avg[dom] * 1e-3) / 1e9
    kernels = {names[c]: {"avg_ms": round(avg[c], 4), "launches": counts[c],
                          "GBps": round(model[c] / (avg[c] * 1e-3) / 1e9, 1)} for c in avg}
    sec_1024 = bench_2d(1024, a.steps, a.warmup) if n != 1024 else None
    mhd = None if a.no_3d else bench_mhd(a.steps, a.warmup)
    if mhd is not None:
        mhd["cpu_baseline"] = cpu3
    ens = None if a.no_ensemble else bench_ensemble(min(a.steps, 20), a.warmup)
    fno = None if a.no_fno else bench_fno(a.fno_epochs)
    fft_flops = 4 * 5 * 2.5 * (n * n) * math.log2(n * n)
    out = {
        "metric": METRIC,
        "value": n * n / (ms_per_step * 1e-3),
        "unit": "grid-pt steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D incompressible hydro, vorticity pseudo-spectral, RK4, {n}^2, FP64",
                   "nx": n, "ny": n, "nu": nu, "dt": dt, "scheme": "rk4", "dealias": "circular 2/3",
                   "ic": "random-phase E(k)~k^4 exp(-(k/4)^2), unit KE, seed 1"},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((names[dom], n, xshape) if dom == 1 else (names[dom], n)),
                     "algorithmic_bytes_per_launch": model[dom],
                     "device_copy_GBps": copy_gbps,
                     "frac_of_copy": round(achieved / copy_gbps, 4) if copy_gbps else None},
        "kernels": kernels,
        "xpass_shape": xshape,
        "step_model_bytes": step_bytes,
        "step_model_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        # FP64 work of the step's transforms: 4 RHS x 5 real 2D transforms of
        # 2.5 N log2 N flops each (N = Nx Ny points), against the vector FP64 peak
        "fp64_fft_model": {"flops_per_step": fft_flops,
                           "tflops": round(fft_flops / (ms_per_step * 1e-3) / 1e12, 2),
                           "peak_tflops": FP64_PEAK_TFLOPS,
                           "frac": round(fft_flops / (ms_per_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)},
        "state_finite": finite,
        "cpu_baseline": cpu,
        "secondary_2d_1024": sec_1024,
        "secondary_3d_mhd": mhd,
        "secondary_ensemble": ens,
        "secondary_fno": fno,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
