"""Continuous batching (GPU): device-to-device state handover, results
independent of the mix (bitwise), HBM budget respected."""

import numpy as np
import pytest

from timemachine_amd import serving
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import custom_ops

pytestmark = pytest.mark.gpu


def _small_job(i, n_waters, W, steps, arrival, heavy=8, barostat=5):
    side = max((n_waters / 33.4) ** (1 / 3), 2.62)
    sys_ = ts.ligand_in_water(side=side, n_waters=n_waters, n_heavy=heavy, n_ligand=int(heavy * 1.5), seed=100 + i)
    return serving.Job(i, sys_, np.linspace(0, 1, W), steps, seed=7 + i, arrival=arrival, barostat_interval=barostat)


@pytest.fixture(scope="module")
def jobs(gpu):
    return [
        _small_job(0, 620, 3, 60, 0.0),
        _small_job(1, 700, 2, 40, 0.0),
        _small_job(2, 640, 4, 80, 0.01),
        _small_job(3, 900, 2, 20, 0.02),
    ]


def test_repack_hands_state_over_bitwise(jobs):
    """Run 30 steps in one batch, or 15 + repack (dropping one system, adding
    another) + 15: the systems that stay end bitwise identical, du/dl series too."""
    job = jobs[0]
    a = serving.make_window_contexts(job)
    ref = custom_ops.BatchedContexts([c for c, _ in a], job.lambdas)
    dl_ref, _, _ = ref.multiple_steps(30, store_du_dl_interval=1)
    x_ref = [ref.get_x_t(s) for s in range(3)]
    v_ref = [ref.get_v_t(s) for s in range(3)]
    box_ref = [ref.get_box(s) for s in range(3)]

    b = serving.make_window_contexts(job)
    extra = serving.make_window_contexts(jobs[1])
    batch = custom_ops.BatchedContexts([c for c, _ in b], job.lambdas)
    dl1, _, _ = batch.multiple_steps(15, store_du_dl_interval=1)
    batch2 = batch.repack([2, 0], [extra[0][0]], [job.lambdas[2], job.lambdas[0], jobs[1].lambdas[0]])
    del batch
    dl2, _, _ = batch2.multiple_steps(15, store_du_dl_interval=1)
    for new_s, old_s in ((0, 2), (1, 0)):
        assert np.array_equal(batch2.get_x_t(new_s), x_ref[old_s])
        assert np.array_equal(batch2.get_v_t(new_s), v_ref[old_s])
        assert np.array_equal(batch2.get_box(new_s), box_ref[old_s])
        assert np.array_equal(np.concatenate([dl1[old_s], dl2[new_s]]), dl_ref[old_s])


def test_repack_validation(jobs):
    a = serving.make_window_contexts(jobs[1])
    batch = custom_ops.BatchedContexts([c for c, _ in a], jobs[1].lambdas)
    with pytest.raises(RuntimeError, match="unique systems"):
        batch.repack([0, 0], [], [0.0, 0.0])
    with pytest.raises(RuntimeError, match="at least one system"):
        batch.repack([], [], [])
    with pytest.raises(RuntimeError, match="one lambda per system"):
        batch.repack([0], [], [0.0, 1.0])


def test_served_mix_equals_each_job_alone(jobs):
    """The budget admits about two jobs at a time, so jobs wait, share the
    engine with different neighbours and leave at different chunk boundaries;
    every job's du/dl series is bitwise the one it gets alone."""
    budget = sum(serving.estimate_system_bytes(j.system.N) * j.windows for j in jobs[:2]) + 1
    srv = serving.Server(budget, max_chunk=30, du_dl_interval=10)
    served = srv.run(jobs)
    alone = serving.run_fifo(jobs, du_dl_interval=10)
    assert [r.job_id for r in served] == [0, 1, 2, 3]
    for r, q in zip(served, alone):
        assert r.du_dl.shape == (jobs[r.job_id].windows, jobs[r.job_id].n_steps // 10)
        assert np.array_equal(r.du_dl, q.du_dl), r.job_id
        assert r.dG == q.dG
        assert r.finished >= r.admitted >= r.arrival
    # admission control: the estimate bounds what the engine holds, within the budget
    for rec in srv.log:
        assert rec["budget_used"] <= budget
        assert rec["device_bytes"] <= rec["budget_used"]
    assert max(rec["systems"] for rec in srv.log) < sum(j.windows for j in jobs)  # never all at once
    stats = serving.latency_stats(served)
    assert stats["p99_s"] >= stats["p50_s"] > 0


def test_minimized_windows_served_equal_fifo(jobs):
    """Jobs whose windows are FIRE-minimised at their lambdas first
    (fe/estimator.py:242-249): the start coordinates differ from the unminimised
    ones, and the served mix still equals FIFO bit for bit."""
    import dataclasses

    mjobs = [dataclasses.replace(j, minimize_steps=20) for j in jobs[:3]]
    wins = serving.make_window_batch(mjobs[0])
    x_min = wins.get_x_t(0)
    assert not np.array_equal(x_min, mjobs[0].system.x)
    assert not np.array_equal(wins.get_x_t(0), wins.get_x_t(len(wins) - 1))  # each window at its own lambda
    budget = sum(serving.estimate_system_bytes(j.system.N) * j.windows for j in mjobs[:2]) + 1
    srv = serving.Server(budget, max_chunk=30, du_dl_interval=10, prepare_ahead=2)  # prepared on a second thread
    served = srv.run(mjobs)
    inline = serving.Server(budget, max_chunk=30, du_dl_interval=10).run(mjobs)
    alone = serving.run_fifo(mjobs, du_dl_interval=10)
    for r, q, z in zip(served, alone, inline):
        assert np.array_equal(r.du_dl, q.du_dl), r.job_id
        assert np.array_equal(z.du_dl, q.du_dl), r.job_id
    assert srv.setup_seconds > 0


def test_window_batch_equals_contexts_bitwise(jobs):
    """BatchedContexts.from_windows (one engine, no Context per window) steps every
    window bitwise as the batch of the W Contexts built from the same arguments."""
    job = jobs[2]
    ctxs = serving.make_window_contexts(job)
    ref = custom_ops.BatchedContexts([c for c, _ in ctxs], job.lambdas)
    dl_ref, fr_ref, bx_ref = ref.multiple_steps(20, store_du_dl_interval=1, store_x_interval=10)
    win = serving.make_window_batch(job)
    dl, fr, bx = win.multiple_steps(20, store_du_dl_interval=1, store_x_interval=10)
    np.testing.assert_array_equal(dl, dl_ref)
    np.testing.assert_array_equal(bx, bx_ref)
    for s in range(job.windows):
        np.testing.assert_array_equal(fr[s], fr_ref[s])
        np.testing.assert_array_equal(win.get_v_t(s), ref.get_v_t(s))
    with pytest.raises(RuntimeError, match="one integrator seed per window"):
        sys_ = job.system
        custom_ops.BatchedContexts.from_windows(sys_.x, np.zeros((2, sys_.N, 3)), sys_.box, None, [1], [], [0.0, 1.0])


def test_gather_hands_state_over_bitwise(jobs):
    """gather takes systems from two batches (device to device) and continues them
    bitwise; a source system may be taken only once."""
    a = serving.make_window_batch(jobs[0])
    b = serving.make_window_batch(jobs[1])
    ra = serving.make_window_batch(jobs[0])
    rb = serving.make_window_batch(jobs[1])
    a.multiple_steps(10, store_du_dl_interval=1)
    b.multiple_steps(10, store_du_dl_interval=1)
    ra.multiple_steps(10, store_du_dl_interval=1)
    rb.multiple_steps(10, store_du_dl_interval=1)
    dl_a, _, _ = ra.multiple_steps(15, store_du_dl_interval=1)
    dl_b, _, _ = rb.multiple_steps(15, store_du_dl_interval=1)
    g = custom_ops.BatchedContexts.gather([(b, 1), (a, 2), (a, 0)],
                                          [jobs[1].lambdas[1], jobs[0].lambdas[2], jobs[0].lambdas[0]])
    del a, b
    dl, _, _ = g.multiple_steps(15, store_du_dl_interval=1)
    np.testing.assert_array_equal(dl[0], dl_b[1])
    np.testing.assert_array_equal(dl[1], dl_a[2])
    np.testing.assert_array_equal(dl[2], dl_a[0])
    np.testing.assert_array_equal(g.get_x_t(1), ra.get_x_t(2))
    np.testing.assert_array_equal(g.get_box(0), rb.get_box(1))
    with pytest.raises(RuntimeError, match="each source system may be taken once"):
        custom_ops.BatchedContexts.gather([(ra, 0), (ra, 0)], [0.0, 0.0])
    with pytest.raises(RuntimeError, match="out of range"):
        custom_ops.BatchedContexts.gather([(ra, 7)], [0.0])


@pytest.mark.slow
def test_served_largest_job_beside_smallest_equals_alone(gpu):
    """The serving workload's extremes in one engine: a 48-window job of ~36k atoms
    (7.1 nm box) beside an 8-window job of ~1.8k atoms (2.62 nm box, the smallest
    box that holds twice cutoff plus padding). 30 steps served (chunks of 10,
    barostat every 15) equal each job run alone, bit for bit."""
    big = _small_job(10, 12000, 48, 30, 0.0, heavy=60, barostat=15)
    small = _small_job(11, 580, 8, 30, 0.0, heavy=20, barostat=15)
    print(f"big job N={big.system.N} x {big.windows}, small job N={small.system.N} x {small.windows}", flush=True)
    assert big.system.N >= 35000 and small.system.N <= 1900
    srv = serving.Server(64 * 2**30, max_chunk=10, du_dl_interval=10)
    served = srv.run([big, small])
    alone = serving.run_fifo([big, small], du_dl_interval=10)
    assert max(rec["systems"] for rec in srv.log) == 56  # both jobs resident together
    for r, q in zip(served, alone):
        np.testing.assert_array_equal(r.du_dl, q.du_dl)
        assert np.all(np.isfinite(r.du_dl))


def test_faulted_job_is_evicted_alone(jobs):
    """A job whose box is smaller than twice cutoff plus padding (2.4 nm) is served
    beside healthy jobs: its windows are flagged on the device (the pair kernel
    skips them), the job leaves with the reference's box error, and every other
    job's du/dlambda series is bitwise the one it gets alone. A job that fails
    while it is prepared (FIRE on the bad box raises) fails alone as well."""
    import dataclasses

    side = 2.4
    sys_bad = ts.ligand_in_water(side=side, n_waters=440, n_heavy=8, n_ligand=12, seed=77)
    assert sys_bad.box[0, 0] < 2 * (1.2 + 0.1)
    bad = serving.Job(9, sys_bad, np.linspace(0, 1, 3), 40, seed=99, arrival=0.0, barostat_interval=5)
    mix = [jobs[0], bad, jobs[1], jobs[3]]
    srv = serving.Server(64 * 2**30, max_chunk=20, du_dl_interval=10)
    served = {r.job_id: r for r in srv.run(mix)}
    alone = {r.job_id: r for r in serving.run_fifo(mix, du_dl_interval=10)}
    assert served[9].du_dl is None and "Cutoff with padding" in served[9].error
    assert alone[9].du_dl is None and "Cutoff with padding" in alone[9].error
    for j in (0, 1, 3):
        assert served[j].error is None
        np.testing.assert_array_equal(served[j].du_dl, alone[j].du_dl)
    assert serving.latency_stats(list(served.values()))["failed"] == 1
    # failing during preparation (FIRE raises on the bad box)
    bad_min = dataclasses.replace(bad, minimize_steps=5)
    res = {r.job_id: r for r in serving.Server(64 * 2**30, max_chunk=20, du_dl_interval=10).run([bad_min, jobs[3]])}
    assert res[9].du_dl is None and "Cutoff with padding" in res[9].error
    np.testing.assert_array_equal(res[3].du_dl, alone[3].du_dl)


def test_batch_status_reports_bad_box_without_raising(gpu):
    """set_raise_faults(False): multiple_steps returns, status() flags only the
    system with the bad box; with the default policy the batch raises."""
    good = ts.ligand_in_water(side=3.0, n_waters=800, n_heavy=12, n_ligand=22, seed=3)
    bad = ts.ligand_in_water(side=2.4, n_waters=440, n_heavy=8, n_ligand=12, seed=77)
    jg = serving.Job(0, good, np.array([0.2, 0.7]), 20, seed=1, barostat_interval=0)
    jb = serving.Job(1, bad, np.array([0.5]), 20, seed=2, barostat_interval=0)
    a, b = serving.make_window_batch(jg), serving.make_window_batch(jb)
    g = custom_ops.BatchedContexts.gather([(a, 0), (b, 0), (a, 1)], [0.2, 0.5, 0.7])
    g.set_raise_faults(False)
    g.multiple_steps(10, store_du_dl_interval=5)
    st = g.status()
    assert st[1] & custom_ops.STATUS_BOX_TOO_SMALL and st[0] == 0 and st[2] == 0
    g.set_raise_faults(True)
    with pytest.raises(RuntimeError, match="Cutoff with padding is more than half of the box width"):
        g.multiple_steps(5, store_du_dl_interval=5)
