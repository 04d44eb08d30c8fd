"""Benchmark: batched lambda-window MD stepping on one MI355X.

Workload (one "step" = one MD step of every resident system): 48 evenly spaced
lambda windows x 4 replicas of the ligand-in-water system (192 systems of
6372 atoms: 2100 flexible waters + a 72-atom synthetic ligand, 4.0 nm box,
rc 1.2 nm, beta 2, 4D decoupling offsets on the ligand), BAOAB at 2.5 fs with
HMR, 300 K, friction 1/ps, MC barostat at 1 bar every 15 steps, du/dlambda
stored every step, all in one engine whose per-step launches are captured into
hipGraphs. value = total atom-steps/s of the batch. The same system stepped
alone through a Context (one system at a time) is reported beside it, and a
serving leg reports p50/p99 job latency of a continuously batched stream of
uneven jobs against one job at a time (FIFO).

Prints ONE JSON line. Only the cpu_baseline leg uses the oracle (FP64 port of
the reference's potentials and BAOAB, timed on the host cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "atom-steps/s (batched λ-windows vs single system); p50/p99 job latency"
PEAK_FP64_TFLOPS = 78.6  # MI355X published FP64 vector peak
# Algorithmic work per unique pair inside the cutoff, du_dx only, counted from
# the reference's FP32 pair code (FMA = 2; rsqrt / exp / rcp = 1): 20 for the
# displacement, minimum image and d^2 (k_nonbonded.cuh:426-434), 27 for the
# electrostatic force with the A&S erfc (nonbonded_common.cuh:38-48, 91-97),
# 9 for the force projection and fixed-point scaling (k_nonbonded.cuh:488-496);
# + 25 when both epsilons are nonzero (Lennard-Jones, nonbonded_common.cuh:
# 121-139); + 3 for the 4D displacement of pairs with a nonzero w
# (k_nonbonded.cuh:437-442).
FLOP_PAIR_BASE = 56
FLOP_PAIR_LJ = 25
FLOP_PAIR_ALCH = 3
# The same count for this build's FP64 pair code (nb_pair.h, with its Newton
# steps and the exp remainder): 71 per pair, + 13 for Lennard-Jones. Reported
# beside the algorithmic figure, not used for the roofline.
IMPL_FLOP_PAIR_BASE = 71
IMPL_FLOP_PAIR_LJ = 13
# HBM bytes per launch of the du_dx tile kernel on this workload: FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0 (factors from a 256 MiB device copy), measured with
# rocprofv3 --pmc on MI355X (summarised under profiles/ with the commands; the
# counter program profiles the boxes as built after FIRE, 0.1% fewer tiles than
# the relaxed state of the kernel trace).
TILE_TRAFFIC_BYTES = 1.0172e9 + 2.1028e9
# Serving admits jobs while the resident atoms stay under this: the batched
# engine's step time is ~0.12 ms + atoms / 3.6e8 atom-steps/s on one MI355X,
# so 800k resident atoms run at ~95% of the peak rate, and admitting more only
# makes every resident job finish later.
SERVING_MAX_RESIDENT_ATOMS = 800_000
# The HIP-event kernel profile (roofline and kernel_ms) runs at one fixed
# setting, the one of the committed rocprofv3 summaries, whatever the autotuner
# chose for the timed steps.
PROFILE_GRID = 65536
PROFILE_PADDING_NM = 0.1


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def pair_counts(sys_, x, box, lam: float):
    """Exact pairs within the 4D cutoff of one state: (all, Lennard-Jones, with w != 0).

    Atoms are grouped by (w, LJ); for two groups whose w differ by dw the 4D
    test d^2 + dw^2 < rc^2 is a 3D test at sqrt(rc^2 - dw^2), counted with
    periodic k-d trees (strict inequality). A pair is alchemical when either
    atom has w != 0."""
    from scipy.spatial import cKDTree

    L = np.diag(box)
    xw = np.mod(x, L)
    xw = np.where(xw >= L, 0.0, xw)
    rc = sys_.cutoff
    w = rc * (sys_.lambda_plane_idxs + sys_.lambda_offset_idxs * lam)
    lj = sys_.nb_params[:, 2] != 0
    groups = []
    for wk, lk in sorted(set(zip(w.tolist(), lj.tolist()))):
        idx = np.nonzero((w == wk) & (lj == lk))[0]
        groups.append((wk, lk, len(idx), cKDTree(xw[idx], boxsize=L)))
    total = total_lj = total_alch = 0
    for a in range(len(groups)):
        for b in range(a, len(groups)):
            dw = groups[a][0] - groups[b][0]
            if dw * dw >= rc * rc:
                continue
            r = np.nextafter(np.sqrt(rc * rc - dw * dw), 0.0)
            c = int(groups[a][3].count_neighbors(groups[b][3], r))
            if a == b:
                c = (c - groups[a][2]) // 2  # self pairs and both orders
            total += c
            if groups[a][1] and groups[b][1]:
                total_lj += c
            if groups[a][0] != 0.0 or groups[b][0] != 0.0:
                total_alch += c
    return total, total_lj, total_alch


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sys_, runs: int = 5, steps_per_run: int = 3):
    """FP64 oracle (reference formulas) MD steps on the host cores: one warm-up
    run, then the median of `runs` timed runs of `steps_per_run` steps each."""
    import torch

    from oracle import bonded as ob
    from oracle import integrator as oi
    from oracle import nonbonded as onb

    threads = int(os.environ.get("TM_CPU_THREADS", min(16, os.cpu_count() or 1)))
    torch.set_num_threads(threads)
    ca, cb, cc = oi.langevin_coefficients(300.0, 0.0025, 1.0, sys_.masses)
    cbs = -cb
    rng = np.random.default_rng(0)
    x = sys_.x.copy()
    v = np.zeros_like(x)
    lam = 0.5

    def one_run():
        nonlocal x, v
        t0 = time.perf_counter()
        for _ in range(steps_per_run):
            f = onb.nonbonded_pairs(x, sys_.nb_params, sys_.box, lam, sys_.exclusion_idxs, sys_.scales, sys_.beta,
                                    sys_.cutoff, sys_.lambda_plane_idxs, sys_.lambda_offset_idxs)[0]
            f = f + ob.bonded("bond", x, sys_.bond_params, lam, sys_.bond_idxs)[0]
            f = f + ob.bonded("angle", x, sys_.angle_params, lam, sys_.angle_idxs)[0]
            f = f + ob.bonded("torsion", x, sys_.torsion_params, lam, sys_.torsion_idxs)[0]
            x, v = oi.baoab_step(x, v, f, 0.0025, ca, cbs, cc, rng.normal(size=x.shape))
        return time.perf_counter() - t0

    one_run()  # warm-up
    times = sorted(one_run() for _ in range(runs))
    med = times[len(times) // 2]
    return {
        "value": sys_.N * steps_per_run / med,
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "runs_s": times,
        "sample": f"median of {runs} runs (after 1 warm-up run) of {steps_per_run} FP64 BAOAB steps (pair-list "
        f"nonbonded + bonded, Torch CPU, {threads} threads) of the {sys_.N}-atom ligand-in-water system at "
        "lambda 0.5",
    }


def serving_jobs(n_jobs: int, step_scale: int, seed: int = 5):
    """The serving stream's jobs (built on the host in worker processes, before
    this process touches the GPU)."""
    from timemachine_amd import serving

    steps = (2000 // step_scale, 20000 // step_scale)
    cfg = dict(atoms=(1800, 36000), windows=(8, 48), steps=steps, heavy=(20, 60), minimize_steps=99)
    workers = min(16, os.cpu_count() or 1)
    log(f"building {n_jobs} serving jobs on the host ({workers} processes)")
    return serving.synthetic_stream(n_jobs, rate_per_s=0.0, seed=seed, workers=workers, **cfg), cfg


def small_serving_jobs(n_jobs: int, seed: int = 9):
    """A second stream of small jobs (1.8k-4k atoms, 4-12 windows, 400-1000
    steps): each job alone leaves most of the GPU idle, which is where
    continuous batching pays."""
    from timemachine_amd import serving

    cfg = dict(atoms=(1800, 4000), windows=(4, 12), steps=(400, 1000), heavy=(20, 30), minimize_steps=99)
    workers = min(16, os.cpu_count() or 1)
    log(f"building {n_jobs} small serving jobs on the host ({workers} processes)")
    return serving.synthetic_stream(n_jobs, rate_per_s=0.0, seed=seed, workers=workers, **cfg), cfg


def serving_leg(jobs, cfg, step_scale: int, seed: int = 5) -> dict:
    """Continuous batching of an uneven job stream vs one job at a time (FIFO).

    The serving workload's distributions: ligand-in-water jobs with atoms
    log-uniform in [1800, 36000] (the lower end is 1800, not 1500, because a
    box must be at least twice cutoff plus padding, 2.62 nm, and that box holds
    ~1.8k atoms), ligands of 20-60 heavy atoms, U{8..48} windows and
    U{2000..20000} steps divided by ``step_scale`` (every window FIRE-minimised
    for 99 steps first, as the reference's simulate does). Capacity is measured
    with all jobs queued at t = 0; the timed stream then arrives as a Poisson
    process at 80% of it, and the same arrivals are replayed through FIFO."""
    from timemachine_amd import serving

    n_jobs = len(jobs)
    work = sum(j.atom_steps for j in jobs)
    fire_work = sum(j.system.N * j.windows * j.minimize_steps for j in jobs)
    budget = 32 * 2**30
    log(f"serving: {n_jobs} jobs, {work:.3g} atom-steps; saturated run for capacity")
    srv = serving.Server(budget, max_chunk=600, du_dl_interval=10, max_resident_atoms=SERVING_MAX_RESIDENT_ATOMS)
    t0 = time.perf_counter()
    srv.run(jobs)
    capacity = work / (time.perf_counter() - t0)
    rate = 0.8 * capacity / (work / n_jobs)
    rng = np.random.default_rng(seed + 1)
    t = 0.0
    for j in jobs:
        t += rng.exponential(1.0 / rate)
        j.arrival = t
    log(f"serving: Poisson arrivals at {rate:.3g} jobs/s (80% of capacity {capacity:.3g} atom-steps/s)")
    srv = serving.Server(budget, max_chunk=600, du_dl_interval=10, max_resident_atoms=SERVING_MAX_RESIDENT_ATOMS)
    t0 = time.perf_counter()
    served = srv.run(jobs)
    span = time.perf_counter() - t0
    log("serving: FIFO one job at a time on the same arrivals")
    t0 = time.perf_counter()
    fifo = serving.run_fifo(jobs, du_dl_interval=10)
    span_fifo = time.perf_counter() - t0
    same = all(np.array_equal(a.du_dl, b.du_dl) for a, b in zip(served, fifo))
    st, sf = serving.latency_stats(served), serving.latency_stats(fifo)
    return {
        "stream": f"{n_jobs} ligand-in-water jobs, atoms log-uniform {cfg['atoms']}, windows U{cfg['windows']}, "
        f"steps U{cfg['steps']} (the serving workload's U(2000, 20000) divided by {step_scale}), "
        f"{cfg['minimize_steps']} FIRE steps per window first, Poisson arrivals at 80% of measured capacity",
        "step_scale": step_scale,
        "job_atoms": [int(j.system.N) for j in jobs],
        "job_windows": [int(j.windows) for j in jobs],
        "jobs": n_jobs,
        "atom_steps": work,
        "fire_atom_steps": fire_work,
        "max_resident_atoms": SERVING_MAX_RESIDENT_ATOMS,
        "capacity_atom_steps_per_s": capacity,
        "arrival_rate_per_s": rate,
        "p50_latency_s": st["p50_s"],
        "p99_latency_s": st["p99_s"],
        "atom_steps_per_s": work / span,
        "max_resident_systems": max(r["systems"] for r in srv.log),
        "chunks": len(srv.log),
        "chunk_seconds": sum(r["seconds"] for r in srv.log),
        "chunk_atom_steps_per_s": work / sum(r["seconds"] for r in srv.log),
        "repack_seconds": srv.repack_seconds,
        "setup_seconds": srv.setup_seconds,
        "fifo": {"p50_latency_s": sf["p50_s"], "p99_latency_s": sf["p99_s"], "atom_steps_per_s": work / span_fifo},
        "results_bitwise_equal_to_fifo": bool(same),
    }


def differentiation_leg(sys_, n_windows: int = 16, n_frames: int = 500, frame_interval: int = 10,
                        fd_frames: int = 10, seed: int = 2024) -> dict:
    """Gradient of the TI loss L = (sum_w c_w <du/dlambda>_w - t)^2, t ~ N(0, 5),
    w.r.t. the ligand's charges, sigmas, epsilons and torsion force constants,
    through the torch.autograd.Function over the du_dp and d(du/dlambda)/dp
    kernels, frames recomputed in the backward pass (not stored). Then central
    differences (h = 1e-4) of the reweighted window averages on the first
    ``fd_frames`` observed frames of every window, against the same estimator
    on those frames, for six parameters spread over the blocks."""
    import torch

    from timemachine_amd.fe import window_gradients as wg
    from timemachine_amd.lib import potentials

    lams = np.linspace(0.0, 1.0, n_windows)
    target = float(np.random.default_rng(seed).normal(0.0, 5.0))
    windows = wg.LambdaWindows(sys_, lams, n_frames=n_frames, frame_interval=frame_interval, seed=seed,
                               minimize_steps=99)
    nb = torch.tensor(sys_.nb_params, dtype=torch.float64, requires_grad=True)
    tk = torch.tensor(sys_.torsion_params[:, 0], dtype=torch.float64, requires_grad=True)
    log(f"differentiation: {n_windows} windows x {n_frames} frames (every {frame_interval} steps), recompute")
    t0 = time.perf_counter()
    loss = wg.ti_loss(nb, tk, windows, target, recompute=True)
    t_fwd = time.perf_counter() - t0
    loss.backward()
    t_total = time.perf_counter() - t0
    lig = sys_.ligand_idxs
    g = nb.grad.numpy()
    grads = {"q": g[lig, 0], "sig": g[lig, 1], "eps": g[lig, 2], "torsion_k": tk.grad.numpy()}

    # the same loss through the trajectory adjoint (pathwise derivative of the
    # same runs: reverse mode through every step, one Hessian-vector product per
    # step, frames recomputed from checkpoints kept by the forward pass)
    nb_a = torch.tensor(sys_.nb_params, dtype=torch.float64, requires_grad=True)
    tk_a = torch.tensor(sys_.torsion_params[:, 0], dtype=torch.float64, requires_grad=True)
    log(f"differentiation: trajectory adjoint, checkpoints every {windows.checkpoint_interval()} steps")
    t0 = time.perf_counter()
    loss_a = wg.ti_loss(nb_a, tk_a, windows, target, recompute=False, method="adjoint")
    t_fwd_a = time.perf_counter() - t0
    loss_a.backward()
    t_total_a = time.perf_counter() - t0
    ga = nb_a.grad.numpy()
    grads_a = {"q": ga[lig, 0], "sig": ga[lig, 1], "eps": ga[lig, 2], "torsion_k": tk_a.grad.numpy()}
    flat_c = np.concatenate([v.ravel() for v in grads.values()])
    flat_a = np.concatenate([v.ravel() for v in grads_a.values()])
    adjoint = {
        "loss": float(loss_a.detach()),
        "loss_equal_to_covariance_run": bool(float(loss_a.detach()) == float(loss.detach())),
        "forward_s": t_fwd_a,
        "forward_backward_s": t_total_a,
        "steps": windows.equil_steps + windows.n_frames * windows.frame_interval,
        "checkpoint_interval": windows.checkpoint_interval(),
        "gradient_norms": {k: float(np.linalg.norm(v)) for k, v in grads_a.items()},
        "cosine_to_covariance_gradient": float(np.dot(flat_a, flat_c) /
                                               max(np.linalg.norm(flat_a) * np.linalg.norm(flat_c), 1e-300)),
    }
    log(f"  adjoint: forward {t_fwd_a:.2f} s, forward + backward {t_total_a:.2f} s, "
        f"cosine to the covariance gradient {adjoint['cosine_to_covariance_gradient']:.3f}")
    # truncated sweep: each stored value's sensitivity followed back 100 steps (0.25 ps)
    windows.adjoint_horizon = 100
    nb_t = torch.tensor(sys_.nb_params, dtype=torch.float64, requires_grad=True)
    tk_t = torch.tensor(sys_.torsion_params[:, 0], dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    wg.ti_loss(nb_t, tk_t, windows, target, recompute=False, method="adjoint").backward()
    t_trunc = time.perf_counter() - t0
    windows.adjoint_horizon = 0
    gt = nb_t.grad.numpy()
    grads_t = {"q": gt[lig, 0], "sig": gt[lig, 1], "eps": gt[lig, 2], "torsion_k": tk_t.grad.numpy()}
    flat_t = np.concatenate([v.ravel() for v in grads_t.values()])
    adjoint["truncated"] = {
        "horizon_steps": 100,
        "forward_backward_s": t_trunc,
        "gradient_norms": {k: float(np.linalg.norm(v)) for k, v in grads_t.items()},
        "cosine_to_covariance_gradient": float(np.dot(flat_t, flat_c) /
                                               max(np.linalg.norm(flat_t) * np.linalg.norm(flat_c), 1e-300)),
    }
    log(f"  adjoint truncated to 100 steps: {t_trunc:.2f} s, cosine to the covariance gradient "
        f"{adjoint['truncated']['cosine_to_covariance_gradient']:.3f}")

    # finite differences on the first fd_frames observed frames of every window
    fdw = wg.LambdaWindows(sys_, lams, n_frames=fd_frames, frame_interval=frame_interval, seed=seed,
                           minimize_steps=99)
    batch, _ = fdw.batch()
    batch.set_grad_observer(0)
    frames = [[] for _ in lams]
    for _ in range(fd_frames):
        batch.multiple_steps(frame_interval - 1)
        for w in range(n_windows):
            frames[w].append((batch.get_x_t(w), batch.get_box(w)))
        batch.observe()
        batch.multiple_steps(1)
    stats = [batch.grad_stats(w) for w in range(n_windows)]
    del batch
    c = wg.trapezoid_weights(lams)
    analytic = fdw.layout.split(sum(c[w] * wg.window_gradient(stats[w], fdw.kT) for w in range(n_windows)))
    nb_impl = sys_.nonbonded().unbound_impl(np.float32)
    tor_impl = potentials.PeriodicTorsion(sys_.torsion_idxs).unbound_impl(np.float32)

    def ti(nb_params, torsion_params, u_ref):
        total, us = 0.0, []
        for w, lam in enumerate(lams):
            a_f, u_f = [], []
            for x, box in frames[w]:
                _, _, dl, u = nb_impl.execute_selective(x, nb_params, box, lam, False, False, True, True)
                _, _, _, ut = tor_impl.execute_selective(x, torsion_params, box, lam, False, False, False, True)
                a_f.append(dl)
                u_f.append(u + ut)
            u_f = np.asarray(u_f)
            us.append(u_f)
            if u_ref is not None:
                logw = -(u_f - u_ref[w]) / fdw.kT
                wts = np.exp(logw - logw.max())
                total += c[w] * float(np.sum(np.asarray(a_f) * wts) / np.sum(wts))
        return total, us

    _, u0 = ti(sys_.nb_params, sys_.torsion_params, None)
    h = 1e-4
    picks = [("q", int(lig[0]), 0), ("q", int(lig[5]), 0), ("sig", int(lig[3]), 1), ("sig", int(lig[10]), 1),
             ("eps", int(lig[7]), 2), ("torsion_k", 2, None)]
    checks = []
    for name, i, col in picks:
        pp, pm = sys_.nb_params.copy(), sys_.nb_params.copy()
        tp, tm = sys_.torsion_params.copy(), sys_.torsion_params.copy()
        if col is None:
            tp[i, 0] += h
            tm[i, 0] -= h
            ana = float(analytic["torsion"][i, 0])
        else:
            pp[i, col] += h
            pm[i, col] -= h
            ana = float(analytic["nonbonded"][i, col])
        fp, _ = ti(pp, tp, u0)
        fm, _ = ti(pm, tm, u0)
        fd = (fp - fm) / (2 * h)
        checks.append({"param": f"{name}[{i}]", "analytic": ana, "central_fd": fd,
                       "rel_err": abs(ana - fd) / max(abs(fd), 1e-12)})
        log(f"  d(TI)/d{name}[{i}]: estimator {ana:.6e}, central difference {fd:.6e}")
    return {
        "workload": f"{n_windows} lambda windows x {n_frames} frames (every {frame_interval} steps, NVT) of the "
        f"{sys_.N}-atom ligand system; L = (TI - t)^2, t = {target:.4f}",
        "loss": float(loss.detach()),
        "forward_s": t_fwd,
        "forward_backward_s": t_total,
        "device_bytes": windows.last_device_bytes,
        "gradient_norms": {k: float(np.linalg.norm(v)) for k, v in grads.items()},
        "n_params": {k: int(v.size) for k, v in grads.items()},
        "fd_frames_per_window": fd_frames,
        "fd_h": h,
        "fd_checks": checks,
        "fd_max_rel_err": max(ch["rel_err"] for ch in checks),
        "adjoint": adjoint,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=600,
                    help="untimed batched steps first: the barostat expands the freshly built boxes by ~8%% over "
                    "the first ~500 steps, and the step time settles with them")
    ap.add_argument("--windows", type=int, default=48)
    ap.add_argument("--replicas", type=int, default=4)
    ap.add_argument("--single-steps", type=int, default=300)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=16, help="jobs in the serving leg (0 skips it)")
    ap.add_argument("--serving-step-scale", type=int, default=5,
                    help="the serving jobs' U(2000, 20000) steps are divided by this")
    ap.add_argument("--diff-frames", type=int, default=500,
                    help="frames per window of the differentiation leg (16 windows; 0 skips it)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--tile-grid", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="skip the occupancy / grid / padding autotuner")
    ap.add_argument("--profile-only", type=int, default=0, metavar="REPS",
                    help="build the batch, warm up, launch each kernel group REPS times eagerly and exit "
                    "(for rocprofv3 counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 256 MiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--adjoint-only", type=int, default=0, metavar="STEPS",
                    help="trajectory adjoint of 16 lambda windows of the ligand system over STEPS steps "
                    "(forward with checkpoints, backward), print its timing and exit (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (lambda windows shard across GPUs "
            "with no per-step exchange; run one process per GPU instead)",
            file=sys.stderr,
        )
        return 2

    if args.copy_calibration:
        import torch

        n = 256 * 2**20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {n} bytes ({n // 1024} KiB read, {n // 1024} KiB written)")
        return 0

    if args.adjoint_only:
        from timemachine_amd import testsystems as ts
        from timemachine_amd.fe import window_gradients as wg

        sys_a = ts.ligand_in_water()
        n = args.adjoint_only - args.adjoint_only % 10
        windows = wg.LambdaWindows(sys_a, np.linspace(0.0, 1.0, 16), n_frames=max(n // 10, 1), frame_interval=10,
                                   seed=2024, minimize_steps=99)
        t0 = time.perf_counter()
        batch, du_dl = windows.adjoint_run()
        t1 = time.perf_counter()
        windows.adjoint_gradients(batch, np.ones(16))
        t2 = time.perf_counter()
        steps = windows.n_frames * windows.frame_interval
        log(f"adjoint of 16 x {sys_a.N} atoms over {steps} steps: forward {(t1 - t0) / steps * 1e3:.3f} ms/step, "
            f"backward {(t2 - t1) / steps * 1e3:.3f} ms/step")
        return 0

    stream = small_stream = None
    if args.serving_jobs > 0 and not args.profile_only:
        stream = serving_jobs(args.serving_jobs, args.serving_step_scale)
        small_stream = small_serving_jobs(args.serving_jobs)

    from timemachine_amd import _lib
    from timemachine_amd import testsystems as ts
    from timemachine_amd.fe import lambda_windows as lw

    _lib.load()
    assert _lib.device_count() >= 1, "no HIP device visible"
    sys_ = ts.ligand_in_water()
    lams = np.linspace(0.0, 1.0, args.windows)
    log(f"building {args.windows * args.replicas} lambda windows of {sys_.N} atoms in one engine "
        "(FIRE-minimised at their lambdas)")
    lam_list, seeds = lw.window_layout(lams, args.replicas)
    batch = lw.build_batch(sys_, lams, args.replicas, barostat_interval=15)
    tuned = None
    if args.tile_grid:
        batch.set_tile_grid(args.tile_grid)
    elif not args.no_autotune and not args.profile_only:
        log("autotuning the pair kernel's occupancy build and grid, then the neighborlist padding, "
            "on a device copy of the batch")
        tuned = batch.autotune()
        log(f"autotuned: {tuned['waves']} waves per SIMD, {tuned['grid']} workgroups, padding {tuned['padding']} nm")
        for c in tuned["candidates"]:
            log(f"  [{c['stage']}] waves {c['waves']} grid {c['grid']:6d} padding {c['padding']:.3f} nm: "
                f"{c['ms_per_step']:.3f} ms/step")
    S = len(batch)
    use_graph = not args.no_graph
    log(f"warmup {args.warmup} batched steps")
    batch.multiple_steps(args.warmup, store_du_dl_interval=1, use_graph=use_graph)
    if args.profile_only:
        prof = batch.profile(reps=args.profile_only)
        log(f"kernel ms: {prof}, tiles {batch.total_tiles()}")
        return 0
    log(f"timing {args.steps} batched steps")
    t0 = time.perf_counter()
    du_dl, _, _ = batch.multiple_steps(args.steps, store_du_dl_interval=1, use_graph=use_graph)
    t_batch = time.perf_counter() - t0  # multiple_steps synchronises the device before returning
    assert np.all(np.isfinite(du_dl))
    atom_steps = S * sys_.N * args.steps
    value = atom_steps / t_batch
    dG, dl_means = lw.ti_free_energy(lams, du_dl, args.replicas)

    log(f"profiling kernels (HIP events) at grid {PROFILE_GRID}, padding {PROFILE_PADDING_NM} nm")
    batch.set_tile_grid(PROFILE_GRID)
    batch.set_nblist_padding(PROFILE_PADDING_NM)
    prof = batch.profile(reps=10)
    tiles = batch.total_tiles()

    # one system at a time, the same kernels through the Context API
    x_min = lw.build_batch(sys_, [0.5], 1, seed0=99, barostat_interval=15).get_x_t(0)  # FIRE-minimised start
    single, single_keep = lw.window_context(sys_, 99, barostat_interval=15, x0=x_min)
    single.multiple_steps(np.full(args.warmup, 0.5), store_du_dl_interval=1)
    t0 = time.perf_counter()
    single.multiple_steps(np.full(args.single_steps, 0.5), store_du_dl_interval=1)
    t_single = time.perf_counter() - t0
    single_rate = sys_.N * args.single_steps / t_single

    # roofline of the pair kernel: algorithmic FLOP per launch / HIP-event duration
    log("counting pairs for the roofline (the state the kernel profile ran on)")
    from concurrent.futures import ThreadPoolExecutor

    states = [(batch.get_x_t(s), batch.get_box(s), lam_list[s]) for s in range(S)]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        counts = list(ex.map(lambda st: pair_counts(sys_, st[0], st[1], float(st[2])), states))
    pairs_total = sum(c[0] for c in counts)
    pairs_lj = sum(c[1] for c in counts)
    pairs_alch = sum(c[2] for c in counts)
    flop = float(FLOP_PAIR_BASE * pairs_total + FLOP_PAIR_LJ * pairs_lj + FLOP_PAIR_ALCH * pairs_alch)
    impl_flop = float(IMPL_FLOP_PAIR_BASE * pairs_total + IMPL_FLOP_PAIR_LJ * pairs_lj)
    achieved = flop / (prof["nonbonded"] * 1e-3) / 1e12
    roofline = {
        "bound": "valu",
        "kernel": "k_nonbonded_tiles (du_dx)",
        "achieved": achieved,
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP64_TFLOPS,
        "traffic": TILE_TRAFFIC_BYTES,
        "flop_per_pair": {"base": FLOP_PAIR_BASE, "lj": FLOP_PAIR_LJ, "alchemical": FLOP_PAIR_ALCH},
        "pairs_per_launch": pairs_total,
        "lj_pairs_per_launch": pairs_lj,
        "alchemical_pairs_per_launch": pairs_alch,
        "flop_per_launch": flop,
        "impl_flop_per_launch": impl_flop,
        "avg_launch_ms": prof["nonbonded"],
        "tiles_per_launch": tiles,
        "pair_evaluations_per_launch": tiles * 64 * 64,
    }
    log(f"roofline state: {pairs_total} pairs ({pairs_lj} LJ, {pairs_alch} alchemical), {tiles} tiles, "
        f"{flop:.4g} FLOP per launch in {prof['nonbonded']:.4f} ms")
    diff = None
    if args.diff_frames > 0:
        diff = differentiation_leg(sys_, n_frames=args.diff_frames)
    serve = serve_small = None
    if stream is not None:
        serve = serving_leg(stream[0], stream[1], args.serving_step_scale)
        serve_small = serving_leg(small_stream[0], small_stream[1], 1, seed=11)
        serve_small["stream"] = serve_small["stream"].replace(
            "steps U(400, 1000) (the serving workload's U(2000, 20000) divided by 1)", "steps U(400, 1000)")
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP64 oracle)")
        cpu = cpu_baseline(sys_)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_batch / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched lambda windows: {args.windows} windows x {args.replicas} replicas of the "
            f"{sys_.N}-atom ligand-in-water system (4.0 nm box, 2100 waters + 72-atom ligand)",
            "systems": S,
            "atoms_per_system": sys_.N,
            "dt_ps": 0.0025,
            "barostat_interval": 15,
            "du_dl_interval": 1,
            "hipgraph": use_graph,
            "autotuned": None if tuned is None else {"waves_per_simd": tuned["waves"], "grid": tuned["grid"],
                                                     "padding_nm": tuned["padding"]},
        },
        "ti_free_energy_kj_mol": dG,
        "du_dl_window_means": dl_means.tolist(),
        "autotune": None if tuned is None else {
            "chosen": {k: v for k, v in tuned.items() if k != "candidates"},
            "candidates": tuned["candidates"],
            "spread": max(c["ms_per_step"] for c in tuned["candidates"])
            / min(c["ms_per_step"] for c in tuned["candidates"]) - 1.0,
        },
        "single_system": {
            "value": single_rate,
            "unit": "atom-steps/s",
            "ms_per_step": 1e3 * t_single / args.single_steps,
            "steps": args.single_steps,
        },
        "batched_over_single": value / single_rate,
        "kernel_ms": prof,
        "roofline": roofline,
        "serving": serve,
        "serving_small_jobs": serve_small,
        "differentiation": diff,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
