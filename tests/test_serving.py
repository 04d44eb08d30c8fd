"""Serving helpers on the CPU (no GPU): memory estimate, TI estimate, job stream."""

import numpy as np
import pytest

from timemachine_amd import serving


def test_estimate_system_bytes_grows_and_covers_tiles():
    a, b = serving.estimate_system_bytes(2000), serving.estimate_system_bytes(36000)
    assert 0 < a < b
    nb = -(-36000 // 64)
    assert b >= (nb * (nb + 1) // 2 + nb) * 64 * 4


def test_ti_estimate_trapezoid():
    lam = np.array([0.0, 0.25, 1.0])
    du_dl = np.array([[1.0, 3.0], [4.0, 4.0], [0.0, 2.0]])  # means 2, 4, 1
    assert serving.ti_estimate(lam, du_dl) == pytest.approx(0.25 * 3 + 0.75 * 2.5)


def test_synthetic_stream_shapes():
    jobs = serving.synthetic_stream(3, rate_per_s=2.0, seed=1, atoms=(1800, 2600), windows=(2, 4), steps=(20, 60),
                                    heavy=(8, 10))
    arr = [j.arrival for j in jobs]
    assert arr == sorted(arr) and arr[0] > 0
    for j in jobs:
        assert 1700 <= j.system.N <= 2700
        assert j.system.box[0, 0] >= 2.62
        assert 2 <= j.windows <= 4 and j.n_steps % 10 == 0


def test_latency_stats():
    rs = [serving.JobResult(i, np.zeros((1, 1)), 0.0, 0.0, 0.0, float(i + 1)) for i in range(100)]
    st = serving.latency_stats(rs)
    assert st["p50_s"] == pytest.approx(50.5) and st["p99_s"] == pytest.approx(99.01)


# ---- the serving loop's scheduling, on a stand-in engine with a scripted clock ----


class _Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 0.0)


class _FakeBatch:
    """Stands in for BatchedContexts: systems are (job id, window, atoms); a step
    costs 1e-3 s plus 1e-9 s per resident atom on the shared clock; du/dl of a
    window at its k-th stored step is job id * 1000 + window * 100 + k."""

    clock = None

    def __init__(self, systems, lam):
        self.systems = list(systems)  # [job_id, window, atoms, steps_done]
        self.sizes = [s[2] for s in self.systems]
        self._lam = list(lam)

    def __len__(self):
        return len(self.systems)

    def set_raise_faults(self, flag):
        pass

    def device_bytes(self):
        return 0

    def status(self):
        return np.zeros(len(self.systems), np.int32)

    def fire_minimize(self, n_steps):
        self.clock.t += n_steps * (1e-3 + 1e-9 * sum(self.sizes))

    def max_force(self):
        return np.zeros(len(self.systems))

    def multiple_steps(self, n, store_du_dl_interval, use_graph=True):
        self.clock.t += n * (1e-3 + 1e-9 * sum(self.sizes))
        k = n // store_du_dl_interval
        out = np.zeros((len(self.systems), k))
        for i, s in enumerate(self.systems):
            k0 = s[3] // store_du_dl_interval
            out[i] = s[0] * 1000 + s[1] * 100 + np.arange(k0, k0 + k)
            s[3] += n
        return out, None, None

    def repack(self, keep, new, lam):
        return _FakeBatch([self.systems[i] for i in keep], lam)

    @staticmethod
    def gather(parts, lam):
        return _FakeBatch([b.systems[i] for b, i in parts], lam)


def _job(i, atoms, windows, steps, arrival, minimize=99):
    sys_ = type("S", (), {"N": atoms})()
    return serving.Job(i, sys_, np.linspace(0, 1, windows), steps, seed=i, arrival=arrival, minimize_steps=minimize)


@pytest.fixture()
def fake_engine(monkeypatch):
    clock = _Clock()
    _FakeBatch.clock = clock
    built = []

    def make(job, minimize=True):
        assert not minimize  # the server minimises admitted jobs together
        built.append(job.job_id)
        return _FakeBatch([[job.job_id, w, job.system.N, 0] for w in range(job.windows)], job.lambdas)

    monkeypatch.setattr(serving, "make_window_batch", make)
    monkeypatch.setattr(serving.custom_ops.BatchedContexts, "gather", staticmethod(_FakeBatch.gather))
    return clock, built


def test_server_admits_under_the_atom_cap_and_stops_chunks_at_arrivals(fake_engine):
    clock, _ = fake_engine
    jobs = [_job(0, 100_000, 8, 400, 0.0), _job(1, 60_000, 4, 200, 0.0), _job(2, 20_000, 4, 100, 0.35),
            _job(3, 30_000, 2, 100, 50.0)]
    srv = serving.Server(1 << 40, max_chunk=600, du_dl_interval=10, clock=clock, sleep=clock.sleep,
                         max_resident_atoms=1_000_000)
    res = srv.run(jobs)
    assert [r.job_id for r in res] == [0, 1, 2, 3] and all(r.error is None for r in res)
    # job 1 would take the resident atoms to 1.04e6 > cap: it waits for job 0
    first = srv.log[0]
    assert first["atoms"] == 800_000
    assert all(rec["atoms"] <= 1_000_000 or rec["systems"] == len(jobs[0].lambdas) for rec in srv.log)
    # every window's du/dl series is complete and in order, whatever the chunking
    for job, r in zip(jobs, res):
        k = job.n_steps // 10
        for w in range(job.windows):
            np.testing.assert_array_equal(r.du_dl[w], job.job_id * 1000 + w * 100 + np.arange(k))
    # job 3 arrives long after the rest finished: the server idles, then runs it at once
    assert res[3].admitted == pytest.approx(50.0)
    # chunks are not cut into fixed 60-step pieces: the first chunk runs to job 2's arrival estimate
    assert max(rec["steps"] for rec in srv.log) > 60


def test_server_minimises_jobs_admitted_together_in_one_engine(fake_engine, monkeypatch):
    clock, built = fake_engine
    calls = []
    real = serving.minimize_together

    def spy(jobs, batches):
        calls.append([j.job_id for j in jobs])
        return real(jobs, batches)

    monkeypatch.setattr(serving, "minimize_together", spy)
    jobs = [_job(0, 10_000, 3, 100, 0.0), _job(1, 10_000, 2, 100, 0.0), _job(2, 10_000, 2, 100, 0.0, minimize=50)]
    srv = serving.Server(1 << 40, max_chunk=600, du_dl_interval=10, clock=clock, sleep=clock.sleep)
    res = srv.run(jobs)
    assert calls == [[0, 1, 2]] and built == [0, 1, 2]
    assert all(r.error is None and r.du_dl.shape == (len(j.lambdas), 10) for r, j in zip(res, jobs))


def test_minimize_together_fails_only_the_job_with_large_forces(monkeypatch):
    clock = _Clock()
    _FakeBatch.clock = clock
    monkeypatch.setattr(serving.custom_ops.BatchedContexts, "gather", staticmethod(_FakeBatch.gather))
    jobs = [_job(0, 1000, 2, 10, 0.0), _job(1, 1000, 3, 10, 0.0)]
    batches = [_FakeBatch([[j.job_id, w, 1000, 0] for w in range(j.windows)], j.lambdas) for j in jobs]
    monkeypatch.setattr(_FakeBatch, "max_force", lambda self: np.array([1.0, 1.0, 1.0, 1e9, 1.0]))
    out = serving.minimize_together(jobs, batches)
    assert out[0][1] == [0, 1] and isinstance(out[1], AssertionError)
