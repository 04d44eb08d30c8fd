"""FIRE minimisation on the GPU (md/minimizer.py:35-163 of the reference).

* The device minimiser equals the FIRE oracle driven step by step with the
  same bound potentials' forces (host loop), to 1e-9 nm: both follow the same
  binary16 step-size sequence, the differences are summation order and
  fixed-point conversion rounding.
* Systems minimised in a batch equal the same systems minimised alone, bit
  for bit; inactive systems are not touched.
* Energy and largest force drop; max_force equals the host norm of the total
  force; minimize_4d passes the reference's force check.
"""

import numpy as np
import pytest

from oracle import fire as ofire
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import LangevinIntegrator, custom_ops
from timemachine_amd.md import minimizer

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.fixture(scope="module")
def hot(gpu):
    """A small ligand in water with the ligand pushed 0.06 nm off its place, so
    the starting forces are large."""
    s = ts.ligand_in_water(side=3.2, n_waters=1050, n_heavy=24, n_ligand=40, seed=5)
    x = s.x.copy()
    x[s.ligand_idxs] += np.array([0.06, -0.03, 0.02])
    return s, x


def _impls(s):
    return [p.bound_impl(F32) for p in s.bound_potentials()]


def _force_fn(impls, box):
    def force(x, lam):
        f = np.zeros_like(x)
        for bp in impls:
            du_dx, _, _ = bp.execute(x, box, lam)
            f -= du_dx
        return f

    return force


@pytest.mark.parametrize("lam", [0.0, 1.0])
def test_device_fire_equals_oracle_driven_by_same_forces(hot, lam):
    s, x0 = hot
    impls = _impls(s)
    sched = np.full(60, lam)
    xd = minimizer.fire_minimize(x0, impls, s.box, sched)
    tr = []
    xo = ofire.fire_minimize(x0, _force_fn(impls, s.box), sched, trace=tr)
    dev = np.abs(xd - xo).max()
    moved = np.abs(xo - x0).max()
    print(f"[lam={lam}] max |x_dev - x_oracle| {dev:.3e} nm, max move {moved:.3e} nm, final dt {float(tr[-1]['dt'])}",
          flush=True)
    assert moved > 1e-4
    assert dev <= 1e-9


def test_energy_and_force_drop(hot):
    s, x0 = hot
    impls = _impls(s)
    sched = np.full(200, 0.0)
    x = minimizer.fire_minimize(x0, impls, s.box, sched)
    u0 = sum(bp.execute(x0, s.box, 0.0)[2] for bp in impls)
    u1 = sum(bp.execute(x, s.box, 0.0)[2] for bp in impls)
    f0 = np.linalg.norm(_force_fn(impls, s.box)(x0, 0.0), axis=1).max()
    f1 = np.linalg.norm(_force_fn(impls, s.box)(x, 0.0), axis=1).max()
    print(f"u {u0:.3f} -> {u1:.3f}; max |F| {f0:.1f} -> {f1:.1f}", flush=True)
    assert u1 < u0 and f1 < f0


def _ctx(s, x, seed):
    impls = _impls(s)
    intg = LangevinIntegrator(300.0, 0.0025, 1.0, s.masses, seed).impl()
    return custom_ops.Context(x, np.zeros_like(x), s.box, intg, impls), impls


def test_batch_equals_alone_and_inactive_untouched(hot):
    s, x0 = hot
    lams = [0.0, 0.4, 1.0]
    alone = []
    for lam in lams:
        alone.append(minimizer.fire_minimize(x0, _impls(s), s.box, np.full(40, lam)))
    ctxs, keep = zip(*[_ctx(s, x0, 10 + w) for w in range(3)])
    batch = custom_ops.BatchedContexts(list(ctxs), lams)
    batch.fire_minimize(lambda_schedules=np.repeat(np.array(lams)[:, None], 40, axis=1))
    for w in range(3):
        np.testing.assert_array_equal(batch.get_x_t(w), alone[w])
    # only system 1 active; n_steps at the current lambdas
    ctxs2, keep2 = zip(*[_ctx(s, x0, 20 + w) for w in range(3)])
    b2 = custom_ops.BatchedContexts(list(ctxs2), lams)
    b2.fire_minimize(n_steps=39, active=[False, True, False])
    np.testing.assert_array_equal(b2.get_x_t(0), x0)
    np.testing.assert_array_equal(b2.get_x_t(2), x0)
    np.testing.assert_array_equal(b2.get_x_t(1), alone[1])


def test_max_force_matches_host_norm(hot):
    s, x0 = hot
    ctx, keep = _ctx(s, x0, 3)
    batch = custom_ops.BatchedContexts([ctx], [0.3])
    mf = batch.max_force()[0]
    ref = np.linalg.norm(_force_fn(keep, s.box)(x0, 0.3), axis=1).max()
    assert mf == pytest.approx(ref, rel=1e-12)


def test_minimize_4d_passes_force_check(hot):
    s, x0 = hot
    impls = _impls(s)
    x = minimizer.minimize_4d(x0, impls, s.masses, s.box, fixed_idxs=s.ligand_idxs, n_fire=50, n_anneal=200)
    assert minimizer.check_forces(x, impls, s.box, 0.0) < minimizer.FORCE_LIMIT
    # the heavy (fixed) ligand barely moves compared with the water
    lig = np.abs(x[s.ligand_idxs] - x0[s.ligand_idxs]).max()
    wat = np.abs(np.delete(x - x0, s.ligand_idxs, axis=0)).max()
    print(f"ligand moved {lig:.2e} nm, water {wat:.2e} nm", flush=True)
    assert lig < wat
