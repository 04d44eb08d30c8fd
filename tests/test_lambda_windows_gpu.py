"""Batched lambda windows at full size: 48 windows x 4 replicas of the
6372-atom ligand-in-water system (192 systems, FIRE-minimised at their
lambdas, MC barostat every 15 steps, du/dlambda stored every step) in one
engine, against windows of different replicas stepped alone through the
Context API from the same minimised coordinates and seeds.

At this size a wave of the tile kernel spans tiles of several systems, the
neighborlist build runs one wave per row block (a lone system splits its row
blocks over 8 waves) and the barostat decides 192 moves at once; none of that
may change a bit of any window's trajectory.
"""

import numpy as np
import pytest

from timemachine_amd import testsystems as ts
from timemachine_amd.fe import lambda_windows as lw

pytestmark = pytest.mark.gpu

WINDOWS, REPLICAS, STEPS = 48, 4, 30


@pytest.fixture(scope="module")
def ligand(gpu):
    return ts.ligand_in_water()


def test_192_window_batch_equals_windows_alone_bitwise(ligand):
    lams = np.linspace(0.0, 1.0, WINDOWS)
    batch = lw.build_batch(ligand, lams, REPLICAS, barostat_interval=15)  # FIRE-minimised at each lambda
    assert len(batch) == WINDOWS * REPLICAS and ligand.N == 6372
    starts = {k: batch.get_x_t(k) for k in (0, 61, 130, 191)}
    du_dl, frames, boxes = batch.multiple_steps(STEPS, store_du_dl_interval=1, store_x_interval=15)
    assert du_dl.shape == (WINDOWS * REPLICAS, STEPS) and np.all(np.isfinite(du_dl))
    lam_list, seeds = lw.window_layout(lams, REPLICAS)
    moved = 0
    for k in (0, 61, 130, 191):  # replicas 0, 1, 2, 3; lambda 0, ~0.28, ~0.72, 1
        print(f"window {k}: replica {k // WINDOWS}, lambda {lam_list[k]:.4f}, seed {seeds[k]}", flush=True)
        ctx, keep = lw.window_context(ligand, seeds[k], barostat_interval=15, x0=starts[k])
        dl_alone, xs_alone, box_alone = ctx.multiple_steps(np.full(STEPS, lam_list[k]), store_du_dl_interval=1,
                                                            store_x_interval=15)
        np.testing.assert_array_equal(du_dl[k], dl_alone)
        np.testing.assert_array_equal(frames[k], xs_alone)
        np.testing.assert_array_equal(boxes[k], box_alone)
        np.testing.assert_array_equal(batch.get_x_t(k), ctx.get_x_t())
        np.testing.assert_array_equal(batch.get_v_t(k), ctx.get_v_t())
        np.testing.assert_array_equal(batch.get_box(k), ctx.get_box())
        moved += int(not np.array_equal(ctx.get_box(), ligand.box))
    assert moved >= 1  # barostat moves were accepted and replayed identically
    dG, means = lw.ti_free_energy(lams, du_dl, REPLICAS)
    assert np.isfinite(dG) and means.shape == (WINDOWS,)
    # du/dlambda vanishes identically at lambda 0 (w = 0 for every atom) and at lambda 1
    # (w = cutoff: the ligand is beyond every pair's reach), and not in between
    assert np.all(du_dl[::WINDOWS] == 0.0) and np.all(du_dl[WINDOWS - 1 :: WINDOWS] == 0.0)
    assert np.all(du_dl[WINDOWS // 2 :: WINDOWS] != 0.0)
    assert np.all(batch.status() == 0)
