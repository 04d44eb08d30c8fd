"""Context.multiple_steps_U (context.cu:139-233) on the GPU.

* u[f, k] is the energy of the f-th stored state at lambda_windows[k]: checked
  against the FP64 oracle on the returned frame and box (|dU| <= 1e-6 |U|).
* The extra evaluations do not perturb the trajectory: frames equal those of
  multiple_steps with the same seeds, bit for bit.
* The batched form gives every system bitwise its single-context result.
"""

import numpy as np
import pytest

from oracle import bonded as obonded
from oracle import nonbonded as onb
from timemachine_amd import serving
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import custom_ops

pytestmark = pytest.mark.gpu

WINDOWS = np.array([0.0, 0.3, 0.65, 1.0])


@pytest.fixture(scope="module")
def job(gpu):
    sys_ = ts.ligand_in_water(side=2.9, n_waters=780, n_heavy=12, n_ligand=22, seed=9)
    return serving.Job(0, sys_, np.array([0.3, 0.8]), 20, seed=21, barostat_interval=5)


def _oracle_u(sys_, x, box, lam):
    u = onb.nonbonded_pairs(x, sys_.nb_params, box, lam, sys_.exclusion_idxs, sys_.scales, sys_.beta, sys_.cutoff,
                            sys_.lambda_plane_idxs, sys_.lambda_offset_idxs)[3]
    for kind, idxs, params in (("bond", sys_.bond_idxs, sys_.bond_params), ("angle", sys_.angle_idxs, sys_.angle_params),
                               ("torsion", sys_.torsion_idxs, sys_.torsion_params)):
        u += obonded.bonded(kind, x, params, lam, idxs)[3]
    return u


def test_context_multiple_steps_u_matches_oracle_and_trajectory(job):
    (ctx, keep), = serving.make_window_contexts(serving.Job(0, job.system, job.lambdas[:1], 20, job.seed,
                                                            barostat_interval=5))
    u, xs, boxes = ctx.multiple_steps_U(0.3, 20, WINDOWS, 5, 5)
    assert u.shape == (4, len(WINDOWS)) and xs.shape == (4, job.system.N, 3) and boxes.shape == (4, 3, 3)
    for f in range(4):
        for k, lam in enumerate(WINDOWS):
            ref = _oracle_u(job.system, xs[f], boxes[f], lam)
            assert abs(u[f, k] - ref) <= 1e-6 * abs(ref), (f, k, u[f, k], ref)
    # same seeds, plain multiple_steps: identical frames
    (twin, keep2), = serving.make_window_contexts(serving.Job(0, job.system, job.lambdas[:1], 20, job.seed,
                                                              barostat_interval=5))
    _, xs2, boxes2 = twin.multiple_steps(np.full(20, 0.3), store_du_dl_interval=5, store_x_interval=5)
    assert np.array_equal(xs, xs2) and np.array_equal(boxes, boxes2)


def test_batched_multiple_steps_u_equals_contexts(job):
    wins = serving.make_window_contexts(job)
    batch = custom_ops.BatchedContexts([c for c, _ in wins], job.lambdas)
    u, frames, boxes = batch.multiple_steps_U(20, WINDOWS, 5, 10)
    assert u.shape == (2, 4, len(WINDOWS)) and boxes.shape == (2, 2, 3, 3)
    for s, lam in enumerate(job.lambdas):
        # a fresh context identical to the batch's window s (same seeds)
        wins_one = serving.make_window_contexts(job)
        ctx = wins_one[s][0]
        u1, xs1, b1 = ctx.multiple_steps_U(lam, 20, WINDOWS, 5, 10)
        assert np.array_equal(u[s], u1)
        assert np.array_equal(frames[s], xs1) and np.array_equal(boxes[s], b1)


def test_multiple_steps_u_validation(job):
    (ctx, keep), = serving.make_window_contexts(serving.Job(0, job.system, job.lambdas[:1], 20, job.seed))
    with pytest.raises(RuntimeError, match="store_u_interval"):
        custom_ops.call("tm_context_multiple_steps_U", ctx.handle, 0.1, 10, 1, custom_ops.pd(np.zeros(1)), 0, 5,
                        custom_ops.pd(np.zeros(2)), custom_ops.pd(np.zeros(60000)), custom_ops.pd(np.zeros(18)))
