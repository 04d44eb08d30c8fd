"""Continuous-batching request loop for free-energy jobs on one GPU.

A job is one system (one topology and parameter set) to be sampled at W
lambda windows for n_steps each, with du/dlambda stored every
``du_dl_interval`` steps and a TI estimate at the end. Jobs arrive over time
and have uneven atom counts, window counts and lengths.

All windows of all admitted jobs are systems of ONE batched engine
(BatchedContexts), so every step of every resident window goes through the
same kernel launches. The loop advances the engine in chunks; at each chunk
boundary it

* retires finished jobs (their windows leave the engine),
* admits waiting jobs that have arrived, first come first served, while their
  estimated device memory fits the HBM budget,
* repacks the engine (BatchedContexts.repack): windows that stay keep their
  coordinates, velocities, box, step counters and barostat state on the
  device (device-to-device handover; only the neighborlists are rebuilt),
  new windows are uploaded once.

Chunks end exactly where the next running window finishes (capped by
``max_chunk``), so no window runs past its length. Fixed-point forces and
counter-based noise make every window's trajectory independent of what else
is resident: a job's du/dlambda series is bitwise the same served alone, in
FIFO order, or in any mix.

The reference runs one process per lambda window per GPU through its worker
pool (timemachine/parallel/client.py:122-153); there is no continuous batching
there, so this module has no counterpart beyond the Context API it drives.
"""

from __future__ import annotations

import dataclasses
import math
import time
from concurrent.futures import Future, ThreadPoolExecutor
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from timemachine_amd import testsystems as ts
from timemachine_amd.lib import LangevinIntegrator, MonteCarloBarostat, custom_ops
from timemachine_amd.md.minimizer import FORCE_LIMIT

TILE = 64


@dataclasses.dataclass
class Job:
    job_id: int
    system: ts.System
    lambdas: np.ndarray
    n_steps: int
    seed: int
    arrival: float = 0.0  # seconds after the start of the stream
    temperature: float = 300.0
    dt: float = 0.0025
    friction: float = 1.0
    barostat_interval: int = 15
    # FIRE steps at each window's lambda before its MD (fe/estimator.py:242-249
    # minimises with a schedule of 100 entries, i.e. 99 steps); 0 skips it
    minimize_steps: int = 0

    @property
    def windows(self) -> int:
        return len(self.lambdas)

    @property
    def atom_steps(self) -> int:
        return self.system.N * self.windows * self.n_steps


@dataclasses.dataclass
class JobResult:
    job_id: int
    du_dl: Optional[np.ndarray]  # [W, n_steps // du_dl_interval]; None for a failed job
    dG: float  # TI estimate, trapezoid over the windows of the per-window mean du/dlambda (nan if failed)
    arrival: float
    admitted: float
    finished: float
    error: Optional[str] = None  # why the job failed (its own fault; other jobs are unaffected)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def estimate_system_bytes(n_atoms: int) -> int:
    """Upper bound of the device memory one system takes in the engine: per-atom
    state and sorted copies (~400 B/atom) plus the neighborlist tile buffer
    (capacity nb (nb + 1) / 2 + nb tiles of 64 column indices and a row)."""
    nb = -(-n_atoms // TILE)
    tiles = nb * (nb + 1) // 2 + nb
    return 400 * (nb * TILE) + tiles * (TILE * 4 + 4) + 4096


def ti_estimate(lambdas: Sequence[float], du_dl: np.ndarray) -> float:
    means = np.asarray(du_dl, dtype=np.float64).mean(axis=1)
    lam = np.asarray(lambdas, dtype=np.float64)
    return float(np.sum(0.5 * (means[1:] + means[:-1]) * np.diff(lam)))


class _Window:
    __slots__ = ("job", "w", "remaining", "chunks")

    def __init__(self, job: Job, w: int):
        self.job = job
        self.w = w
        self.remaining = job.n_steps
        self.chunks: List[np.ndarray] = []


def window_seed(job: Job, w: int) -> int:
    """Integrator seed of window w of a job (the barostat's is this + 17)."""
    return job.seed * 1009 + w


def make_window_contexts(job: Job):
    """One Context per lambda window of the job, seeded as make_window_batch
    seeds its windows (not FIRE-minimised): the one-Context-per-window form of
    the same job, for checks against the batch."""
    sys_ = job.system
    out = []
    bps = [p.bound_impl(np.float32) for p in sys_.bound_potentials()]
    for w in range(job.windows):
        seed = window_seed(job, w)
        intg = LangevinIntegrator(job.temperature, job.dt, job.friction, sys_.masses, seed).impl()
        baro = None
        if job.barostat_interval > 0:
            baro = MonteCarloBarostat(sys_.N, 1.0, job.temperature, sys_.groups, job.barostat_interval,
                                      seed + 17).impl(bps)
        v0 = ts.maxwell_boltzmann(sys_.masses, job.temperature, seed)
        ctx = custom_ops.Context(sys_.x, v0, sys_.box, intg, bps, barostat=baro)
        out.append((ctx, (bps, intg, baro)))
    return out


def make_window_batch(job: Job, minimize: bool = True) -> custom_ops.BatchedContexts:
    """All lambda windows of a job in one engine (one constructor call, no
    Context per window), seeded per job and window, and FIRE-minimised at each
    window's lambda on the device when the job asks for it (and ``minimize``)."""
    sys_ = job.system
    bps = [p.bound_impl(np.float32) for p in sys_.bound_potentials()]
    seeds = [window_seed(job, w) for w in range(job.windows)]
    # the integrator's seed is replaced per window; its coefficients are shared
    intg = LangevinIntegrator(job.temperature, job.dt, job.friction, sys_.masses, seeds[0]).impl()
    baro = None
    if job.barostat_interval > 0:
        baro = MonteCarloBarostat(sys_.N, 1.0, job.temperature, sys_.groups, job.barostat_interval,
                                  seeds[0] + 17).impl(bps)
    v0s = np.stack([ts.maxwell_boltzmann(sys_.masses, job.temperature, sd) for sd in seeds])
    batch = custom_ops.BatchedContexts.from_windows(sys_.x, v0s, sys_.box, intg, seeds, bps, job.lambdas,
                                                     barostat=baro, barostat_seeds=[sd + 17 for sd in seeds])
    if minimize and job.minimize_steps > 0:
        batch.fire_minimize(n_steps=job.minimize_steps)
        worst = batch.max_force()
        if not np.all(worst < FORCE_LIMIT):
            raise AssertionError(f"job {job.job_id}: forces much greater than expected after minimization")
    return batch


def minimize_together(jobs: Sequence[Job], batches: Sequence[custom_ops.BatchedContexts]):
    """FIRE-minimises the windows of several jobs in one engine (one launch
    sequence for all of them instead of one per job); jobs with the same
    number of FIRE steps share a run. FIRE is per system (each window's own
    step size and mixing), so every window ends bitwise where its own job's
    batch would. Returns, per job, the batch and the window indices holding
    its minimised windows, or the AssertionError of a job whose forces stayed
    too large, or the RuntimeError of a job with a faulted window (bad box)."""
    out: List = [None] * len(jobs)
    groups: Dict[int, List[int]] = {}
    for k, job in enumerate(jobs):
        if job.minimize_steps > 0:
            groups.setdefault(job.minimize_steps, []).append(k)
        else:
            out[k] = (batches[k], list(range(job.windows)))
    for n_steps, ks in groups.items():
        if len(ks) == 1:
            merged, offs = batches[ks[0]], [0]
        else:
            parts, lam, offs = [], [], []
            for k in ks:
                offs.append(len(parts))
                parts += [(batches[k], w) for w in range(jobs[k].windows)]
                lam += [float(v) for v in jobs[k].lambdas]
            merged = custom_ops.BatchedContexts.gather(parts, lam)
        # a window with a bad box (or another fault) fails its job only: the
        # engine flags it, skips its pairs and keeps minimising the others
        merged.set_raise_faults(False)
        merged.fire_minimize(n_steps=n_steps)
        worst = merged.max_force()
        status = merged.status()
        for k, o in zip(ks, offs):
            W = jobs[k].windows
            bad = np.nonzero(status[o : o + W])[0]
            if bad.size:
                out[k] = RuntimeError(f"window {int(bad[0])}: " + custom_ops.status_message(int(status[o + bad[0]])))
            elif np.all(worst[o : o + W] < FORCE_LIMIT):
                out[k] = (merged, list(range(o, o + W)))
            else:
                out[k] = AssertionError(f"job {jobs[k].job_id}: forces much greater than expected after minimization")
    return out


class Server:
    """Continuous batching of jobs into one engine under an HBM budget.

    Admission is first come first served while the job fits the HBM budget
    and, when ``max_resident_atoms`` is set, while the resident atoms stay
    under it (the engine's throughput saturates at about a million atoms on
    one MI355X; admitting beyond that only slows every resident job). A job
    is always admitted onto an idle GPU. A chunk runs until the next event
    the loop has to act on: a window finishing, or the next arrival while
    that job would fit, estimated from the last chunk's step time; at most
    ``max_chunk`` steps."""

    def __init__(self, hbm_budget_bytes: int, max_chunk: int = 600, du_dl_interval: int = 10,
                 use_graph: bool = True, clock: Callable[[], float] = time.perf_counter,
                 sleep: Callable[[float], None] = time.sleep, prepare_ahead: int = 0,
                 max_resident_atoms: Optional[int] = None):
        if du_dl_interval <= 0 or max_chunk < du_dl_interval:
            raise ValueError("need 0 < du_dl_interval <= max_chunk")
        self.max_resident_atoms = None if max_resident_atoms is None else int(max_resident_atoms)
        self._step_s = None  # seconds per step of the last chunk
        self.budget = int(hbm_budget_bytes)
        self.max_chunk = int(max_chunk) - int(max_chunk) % int(du_dl_interval)
        self.du_dl_interval = int(du_dl_interval)
        self.use_graph = use_graph
        self._clock = clock
        self._sleep = sleep
        self.log: List[dict] = []  # one record per chunk (resident systems, steps, duration)
        self.repack_seconds = 0.0  # time spent building engines at admissions and retirements
        # window contexts of arrived jobs, with their FIRE minimisation: built at
        # admission on the serving thread (prepare_ahead = 0), or on a second host
        # thread (its own engine and HIP stream) while the batch runs, at most
        # prepare_ahead jobs ahead. The second thread overlaps host work only:
        # allocations and frees synchronise the whole device, so its GPU work
        # waits for the running chunk.
        self.prepare_ahead = int(prepare_ahead)
        self.setup_seconds = 0.0  # total preparation time (overlapped with chunks when prepare_ahead > 0)
        self.setup_wait_seconds = 0.0  # time the serving thread waited for a preparation

    def _prepare(self, job: Job):
        """The job's window batch, or the exception that preparing it raised
        (a bad box, forces too large after minimisation): only that job fails."""
        t = time.perf_counter()
        try:
            wins = make_window_batch(job)
            wins.set_raise_faults(False)
        except (RuntimeError, AssertionError) as e:
            wins = e
        self.setup_seconds += time.perf_counter() - t
        return wins

    def _job_bytes(self, job: Job) -> int:
        return job.windows * estimate_system_bytes(job.system.N)

    def _fits(self, job: Job, used: int, resident_atoms: int, busy: bool) -> bool:
        """Whether job can be admitted now: within the HBM budget and, when
        the GPU is busy, within the resident-atom cap."""
        if used + self._job_bytes(job) > self.budget:
            return False
        if busy and self.max_resident_atoms is not None:
            return resident_atoms + job.system.N * job.windows <= self.max_resident_atoms
        return True

    def run(self, jobs: Sequence[Job]) -> List[JobResult]:
        for j in jobs:
            if j.n_steps % self.du_dl_interval != 0:
                raise ValueError(f"job {j.job_id}: n_steps must be a multiple of du_dl_interval")
            if self._job_bytes(j) > self.budget:
                raise ValueError(f"job {j.job_id} alone exceeds the HBM budget")
        pending = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        t0 = self._clock()
        now = lambda: self._clock() - t0  # noqa: E731
        running: List[_Window] = []
        batch: Optional[custom_ops.BatchedContexts] = None
        used = 0
        admitted_at: Dict[int, float] = {}
        results: List[JobResult] = []
        pool = ThreadPoolExecutor(max_workers=1) if self.prepare_ahead > 0 else None
        prepared: Dict[int, Future] = {}
        try:
            while pending or running:
                batch, used = self._serve_once(pending, running, batch, used, admitted_at, results, pool, prepared,
                                               now)
        finally:
            if pool is not None:
                pool.shutdown(wait=True)
        return sorted(results, key=lambda r: r.job_id)

    def _serve_once(self, pending, running, batch, used, admitted_at, results, pool, prepared, now):
        """Admit what fits, run one chunk, retire finished jobs."""
        if pool is not None:
            # prepare arrived jobs in admission order, a few ahead
            for job in pending[: self.prepare_ahead]:
                if job.arrival <= now() and job.job_id not in prepared:
                    prepared[job.job_id] = pool.submit(self._prepare, job)
        # admit, first come first served, while the HBM budget and the atom cap allow
        resident = sum(r.job.system.N for r in running)
        new_parts = []  # (job, batch, that batch's system index of each window)
        to_minimize: List[Job] = []
        unminimized = []
        while pending and pending[0].arrival <= now() and self._fits(
                pending[0], used, resident, bool(running or new_parts or to_minimize)):
            job = pending[0]
            if pool is not None:
                fut = prepared.get(job.job_id)
                if fut is None:
                    fut = prepared[job.job_id] = pool.submit(self._prepare, job)
                if not fut.done() and (running or new_parts):
                    break  # keep the batch busy while the job is prepared
                tw = now()
                wins = fut.result()
                self.setup_wait_seconds += now() - tw
                del prepared[job.job_id]
            else:
                t = time.perf_counter()
                try:
                    wins = make_window_batch(job, minimize=False)
                except (RuntimeError, AssertionError) as e:
                    wins = e
                self.setup_seconds += time.perf_counter() - t
            pending.pop(0)
            if isinstance(wins, Exception):
                t_fail = now()
                results.append(JobResult(job.job_id, None, float("nan"), job.arrival, t_fail, t_fail, error=str(wins)))
                continue
            used += self._job_bytes(job)
            resident += job.system.N * job.windows
            admitted_at[job.job_id] = now()
            if pool is not None:
                new_parts.append((job, wins, list(range(job.windows))))
            else:
                to_minimize.append(job)
                unminimized.append(wins)
        if to_minimize:
            # the windows of every job admitted now are minimised in one engine
            t = time.perf_counter()
            done = minimize_together(to_minimize, unminimized)
            self.setup_seconds += time.perf_counter() - t
            del unminimized
            for job, m in zip(to_minimize, done):
                if isinstance(m, Exception):
                    t_fail = now()
                    results.append(JobResult(job.job_id, None, float("nan"), job.arrival, admitted_at[job.job_id],
                                             t_fail, error=str(m)))
                    used -= self._job_bytes(job)
                    continue
                new_parts.append((job, m[0], m[1]))
        new: List[_Window] = [_Window(job, w) for job, _, _ in new_parts for w in range(job.windows)]
        if not running and not new:
            if pending:
                # idle until the next arrival
                self._sleep(max(0.0, pending[0].arrival - now()))
            return batch, used
        if new:
            tr = now()
            lam = [r.job.lambdas[r.w] for r in running] + [n.job.lambdas[n.w] for n in new]
            parts = [(batch, s) for s in range(len(running))]
            parts += [(b, i) for _, b, idxs in new_parts for i in idxs]
            only = new_parts[0] if len(new_parts) == 1 else None
            if batch is None and only is not None and only[2] == list(range(len(only[1]))):
                batch = only[1]  # the job's own engine becomes the batch
            else:
                batch = custom_ops.BatchedContexts.gather(parts, lam)
            batch.set_raise_faults(False)
            del new_parts
            running += new
            self.repack_seconds += now() - tr
        chunk = min(self.max_chunk, min(r.remaining for r in running))
        if pending and self._fits(pending[0], used, sum(r.job.system.N for r in running), True):
            # stop at the first boundary after the next job arrives (it would be admitted there)
            wait = pending[0].arrival - now()
            if self._step_s is None:
                steps = 60
            else:
                steps = int(math.ceil(max(wait, 0.0) / self._step_s))
            steps = max(self.du_dl_interval, -(-steps // self.du_dl_interval) * self.du_dl_interval)
            chunk = min(chunk, steps)
        tc = now()
        du_dl, _, _ = batch.multiple_steps(chunk, store_du_dl_interval=self.du_dl_interval,
                                           use_graph=self.use_graph)
        dt = now() - tc
        self._step_s = dt / chunk
        self.log.append({"t": tc, "systems": len(running), "atoms": int(sum(batch.sizes)), "steps": chunk,
                         "seconds": dt, "budget_used": used, "device_bytes": batch.device_bytes()})
        for s, r in enumerate(running):
            r.chunks.append(du_dl[s])
            r.remaining -= chunk
        # a job with a faulted window (bad box, non-finite state, fixed-point
        # overflow) leaves with an error; the other jobs are not affected
        status = batch.status()
        failed: Dict[int, str] = {}
        for s, r in enumerate(running):
            if status[s] and r.job.job_id not in failed:
                failed[r.job.job_id] = f"window {r.w}: " + custom_ops.status_message(int(status[s]))
        done_jobs = {r.job.job_id for r in running if r.remaining == 0} | set(failed)
        if done_jobs:
            t_fin = now()
            by_job: Dict[int, List[_Window]] = {}
            for r in running:
                if r.job.job_id in done_jobs:
                    by_job.setdefault(r.job.job_id, []).append(r)
            for jid, ws in by_job.items():
                ws.sort(key=lambda r: r.w)
                job = ws[0].job
                if jid in failed:
                    results.append(JobResult(jid, None, float("nan"), job.arrival, admitted_at[jid], t_fin,
                                             error=failed[jid]))
                else:
                    series = np.stack([np.concatenate(r.chunks) for r in ws])
                    results.append(JobResult(jid, series, ti_estimate(job.lambdas, series), job.arrival,
                                             admitted_at[jid], t_fin))
                used -= self._job_bytes(job)
            keep = [i for i, r in enumerate(running) if r.job.job_id not in done_jobs]
            tr = now()
            if keep:
                lam = [running[i].job.lambdas[running[i].w] for i in keep]
                batch = batch.repack(keep, [], lam)
                batch.set_raise_faults(False)
            else:
                batch = None
            self.repack_seconds += now() - tr
            running[:] = [running[i] for i in keep]
        return batch, used


def run_fifo(jobs: Sequence[Job], du_dl_interval: int = 10, use_graph: bool = True,
             clock: Callable[[], float] = time.perf_counter,
             sleep: Callable[[float], None] = time.sleep) -> List[JobResult]:
    """One job at a time in arrival order: each job's windows run as one batch."""
    t0 = clock()
    now = lambda: clock() - t0  # noqa: E731
    results = []
    for job in sorted(jobs, key=lambda j: (j.arrival, j.job_id)):
        if job.arrival > now():
            sleep(job.arrival - now())
        ta = now()
        try:
            batch = make_window_batch(job)
            batch.set_raise_faults(False)
            du_dl, _, _ = batch.multiple_steps(job.n_steps, store_du_dl_interval=du_dl_interval,
                                               use_graph=use_graph)
            status = batch.status()
            del batch
        except (RuntimeError, AssertionError) as e:
            results.append(JobResult(job.job_id, None, float("nan"), job.arrival, ta, now(), error=str(e)))
            continue
        bad = np.nonzero(status)[0]
        if bad.size:
            msg = f"window {int(bad[0])}: " + custom_ops.status_message(int(status[bad[0]]))
            results.append(JobResult(job.job_id, None, float("nan"), job.arrival, ta, now(), error=msg))
            continue
        results.append(JobResult(job.job_id, du_dl, ti_estimate(job.lambdas, du_dl), job.arrival, ta, now()))
    return results


def _build_job_system(n_target: float, n_heavy: int, seed: int) -> ts.System:
    """A ligand-in-water system of about n_target atoms (the box is at least
    2.62 nm, twice cutoff plus padding), deterministic in its arguments."""
    rng = np.random.default_rng(seed)
    n_waters = max(int((n_target - 1.6 * n_heavy) / 3), 1)
    side = max((n_waters / 33.4) ** (1.0 / 3.0), 2.62)
    n_lig = int(round(n_heavy * 1.6))
    for _ in range(12):
        try:
            return ts.ligand_in_water(side=side, n_waters=n_waters, n_heavy=n_heavy, n_ligand=n_lig,
                                      seed=int(rng.integers(1 << 30)))
        except ValueError:
            n_lig = max(n_heavy, n_lig - max(1, n_heavy // 10))  # no room for every hydrogen
    raise RuntimeError("could not build a synthetic ligand")


def synthetic_stream(n_jobs: int, rate_per_s: float, seed: int = 0, atoms=(1800, 36000), windows=(8, 48),
                     steps=(2000, 20000), heavy=(20, 60), steps_multiple: int = 10,
                     minimize_steps: int = 0, workers: int = 1) -> List[Job]:
    """Poisson arrivals of ligand-in-water jobs: atoms log-uniform, windows and
    steps uniform; each window FIRE-minimised for ``minimize_steps`` steps
    before its MD. The systems are built on the host, in ``workers`` processes
    (started fresh, so call this before the process uses the GPU)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / rate_per_s) if rate_per_s > 0 else 0.0
        n_target = math.exp(rng.uniform(math.log(atoms[0]), math.log(atoms[1])))
        n_heavy = int(rng.integers(heavy[0], heavy[1] + 1))
        sys_seed = int(rng.integers(1 << 30))
        W = int(rng.integers(windows[0], windows[1] + 1))
        n = int(rng.integers(steps[0], steps[1] + 1))
        n -= n % steps_multiple
        specs.append((i, t, n_target, n_heavy, sys_seed, W, max(n, steps_multiple), int(rng.integers(1 << 30))))
    args = [(sp[2], sp[3], sp[4]) for sp in specs]
    if workers > 1:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor

        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            systems = list(ex.map(_build_job_system, *zip(*args)))
    else:
        systems = [_build_job_system(*a) for a in args]
    return [Job(i, sys_, np.linspace(0.0, 1.0, W), n, job_seed, arrival=t, minimize_steps=minimize_steps)
            for (i, t, _, _, _, W, n, job_seed), sys_ in zip(specs, systems)]


def latency_stats(results: Sequence[JobResult]) -> dict:
    """Latency percentiles of the jobs that completed (failed jobs are counted apart)."""
    lat = np.array([r.latency for r in results if r.error is None])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "jobs": len(lat), "failed": sum(r.error is not None for r in results)}
