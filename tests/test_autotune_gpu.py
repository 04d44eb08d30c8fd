"""Occupancy / grid / padding autotuner (GPU): it changes speed only. A batch
that was autotuned (on its device copy) and then stepped gives bitwise the
trajectory of an untuned batch; explicit occupancy builds, padding and grid
settings do not change bits."""

import numpy as np
import pytest

from timemachine_amd import serving
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import custom_ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def job(gpu):
    sys_ = ts.ligand_in_water(side=2.9, n_waters=780, n_heavy=12, n_ligand=22, seed=3)
    return serving.Job(0, sys_, np.linspace(0, 1, 4), 40, seed=5, barostat_interval=10)


def _batch(job):
    wins = serving.make_window_contexts(job)
    return custom_ops.BatchedContexts([c for c, _ in wins], job.lambdas), wins


def test_autotune_keeps_trajectories(job):
    ref, keep_r = _batch(job)
    dl_ref, _, _ = ref.multiple_steps(40, store_du_dl_interval=1)
    tuned, keep_t = _batch(job)
    res = tuned.autotune(steps=6, waves=(4, 2), grids=(2048, 8192), paddings=(0.1, 0.2))
    assert {r["grid"] for r in res["candidates"]} == {2048, 8192}
    assert {r["waves"] for r in res["candidates"]} == {4, 2}
    # the choice is the fastest of the last (padding) stage
    assert res["ms_per_step"] == min(r["ms_per_step"] for r in res["candidates"] if r["stage"] == "padding")
    dl, _, _ = tuned.multiple_steps(40, store_du_dl_interval=1)
    assert np.array_equal(dl, dl_ref)
    for s in range(len(job.lambdas)):
        assert np.array_equal(tuned.get_x_t(s), ref.get_x_t(s))


def test_padding_and_grid_do_not_change_bits(job):
    ref, keep_r = _batch(job)
    dl_ref, _, _ = ref.multiple_steps(25, store_du_dl_interval=1)
    b, keep_b = _batch(job)
    b.set_nblist_padding(0.05)
    b.set_tile_grid(512)
    b.set_tile_waves(2)
    dl1, _, _ = b.multiple_steps(8, store_du_dl_interval=1)
    b.set_nblist_padding(0.2)
    b.set_tile_grid(0)
    b.set_tile_waves(3)
    dl2, _, _ = b.multiple_steps(8, store_du_dl_interval=1)
    b.set_tile_waves(4)
    dl3, _, _ = b.multiple_steps(9, store_du_dl_interval=1)
    assert np.array_equal(np.concatenate([dl1, dl2, dl3], axis=1), dl_ref)
    with pytest.raises(RuntimeError, match="waves per SIMD must be 2, 3 or 4"):
        b.set_tile_waves(5)
    with pytest.raises(RuntimeError, match="padding must be"):
        b.set_nblist_padding(-1.0)
    with pytest.raises(RuntimeError, match="tile grid must be"):
        b.set_tile_grid(-1)
    with pytest.raises(RuntimeError, match="tile grid must be"):
        b.set_tile_grid((1 << 24) + 1)
