"""Autotuner host logic (CPU): candidates are timed on a device copy, twice, in
one order and then in the reverse order, and ranked by their mean; stage 1
ranks the pair kernel's occupancy build x grid at padding 0.1 nm, stage 2 the
paddings at the stage-1 winner. A step time
that drifts while the copy runs (a freshly built system relaxing) then adds
the same amount to every candidate to first order (exactly, for a linear
drift: the two timing windows of every candidate have the same mean
position), instead of favouring whichever candidate is timed last. A stand-in
batch with a scripted clock replaces the device."""

import time

import pytest

from timemachine_amd.lib import custom_ops


class _FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class _FakeBatch:
    """Steps cost cost[(waves, grid, padding)] plus a drift that falls linearly
    with the steps taken, from drift0 by drift_per_step per step."""

    def __init__(self, clock, cost, drift0, drift_per_step):
        self.clock = clock
        self.cost = cost
        self.drift0 = drift0
        self.drift = drift_per_step
        self.steps_done = 0
        self.waves = None
        self.grid = None
        self.pad = None
        self.calls = []
        self._lam = [0.0]

    def __len__(self):
        return 1

    def repack(self, keep, new, lam):
        return self

    def set_tile_waves(self, w):
        self.waves = w

    def set_tile_grid(self, g):
        self.grid = g

    def set_nblist_padding(self, p):
        self.pad = p

    def multiple_steps(self, n, **kw):
        for _ in range(n):
            base = self.cost[(self.waves, self.grid, self.pad)] if self.pad is not None else 1.0
            self.clock.t += base + self.drift0 - self.drift * self.steps_done
            self.steps_done += 1
        self.calls.append((self.waves, self.grid, self.pad, n))
        return None, None, None


@pytest.fixture()
def fake(monkeypatch):
    clock = _FakeClock()
    monkeypatch.setattr(time, "perf_counter", clock)
    # stage 1 (padding 0.1): the last candidate of the first order is slightly
    # worse than the first one, but the drift makes it look best when every
    # candidate is timed once; stage 2 then prefers padding 0.2 for (4, 1)
    cost = {(4, 1, 0.1): 1.00, (4, 2, 0.1): 1.02, (2, 1, 0.1): 1.03, (2, 2, 0.1): 1.01,
            (4, 1, 0.2): 0.98, (4, 1, 0.05): 1.04}
    return _FakeBatch(clock, cost, drift0=0.4, drift_per_step=1e-3)


def test_autotune_ranks_by_the_mean_of_two_opposite_sweeps(fake):
    res = custom_ops.BatchedContexts.autotune(fake, steps=30, waves=(4, 2), grids=(1, 2), paddings=(0.05, 0.1, 0.2))
    assert (res["waves"], res["grid"], res["padding"]) == (4, 1, 0.2)
    assert fake.waves == 4 and fake.grid == 1 and fake.pad == 0.2  # the choice is applied
    timed = [(w, g, p) for w, g, p, n in fake.calls if n == 30][1:]  # after the warm-up call
    stage1 = [(4, 1, 0.1), (4, 2, 0.1), (2, 1, 0.1), (2, 2, 0.1)]
    stage2 = [(4, 1, 0.05), (4, 1, 0.1), (4, 1, 0.2)]
    assert timed == stage1 + stage1[::-1] + stage2 + stage2[::-1]
    assert len(res["candidates"]) == 7
    assert [r["stage"] for r in res["candidates"]] == ["occupancy x grid"] * 4 + ["padding"] * 3


def test_single_sweep_would_pick_the_drifted_candidate(fake):
    """The failure the two sweeps prevent: timed once, in order, the last
    candidate wins only because the drift had fallen by then."""
    fake.set_nblist_padding(0.1)
    fake.set_tile_waves(4)
    fake.set_tile_grid(1)
    fake.multiple_steps(30)
    times = {}
    for w, g in [(4, 1), (4, 2), (2, 1), (2, 2)]:
        fake.set_tile_waves(w)
        fake.set_tile_grid(g)
        t0 = fake.clock()
        fake.multiple_steps(30)
        times[(w, g)] = fake.clock() - t0
    assert min(times, key=times.get) != (4, 1)
