"""Monte Carlo barostat on the device (barostat.cu:139-373).

* the decision: the device's accept flag and new box equal the oracle's
  acceptance rule applied to the energies, volume change and draw the device
  used (barostat.cu:220-279), and its energies equal the potentials'
  fixed-point sums of the same states;
* ideal gas: without nonbonded terms, <V> = (N_mol + 1) kT / P within 1% at
  three temperatures (tests/test_barostat.py:262-365, OpenMM's testIdealGas);
* a subset of the molecules as groups (tests/test_barostat.py:69-127);
* pressure changes (tests/test_barostat.py:196-259);
* the move stream continues across engine rebuilds (batch hand-over,
  set_pressure): a system stepped inside and outside batches equals the same
  system stepped in one Context.
"""

import numpy as np
import pytest

from oracle import barostat as obaro
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import LangevinIntegrator, MonteCarloBarostat, custom_ops

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def ligand(gpu):
    return ts.ligand_in_water()


def _context(sys_, seed, interval, pressure=1.0, temperature=300.0, groups=None, bps=None, box=None, x=None,
             dt=0.0025):
    if bps is None:
        bps = [p.bound_impl(F32) for p in sys_.bound_potentials()]
    intg = LangevinIntegrator(temperature, dt, 1.0, sys_.masses, seed).impl()
    baro = MonteCarloBarostat(sys_.N, pressure, temperature, sys_.groups if groups is None else groups, interval,
                              seed + 1).impl(bps)
    v0 = ts.maxwell_boltzmann(sys_.masses, temperature, seed)
    ctx = custom_ops.Context(sys_.x if x is None else x, v0, sys_.box if box is None else box, intg, bps,
                             barostat=baro)
    return ctx, (bps, intg, baro)


@pytest.mark.parametrize("pressure", [1.0, 2000.0])
def test_move_decision_matches_oracle(ligand, pressure):
    """For several moves: accept == oracle.accept(acceptance_work(...)), the new box
    is the old one scaled by the cube root of the volume ratio, and u_before / u_after
    are the energies of the pre-move state and of its centroid-scaled image."""
    interval = 4
    ctx, keep = _context(ligand, 11, interval, pressure=pressure)[:2]
    batch = custom_ops.BatchedContexts([ctx], [0.3])
    bps = keep[0]
    nmol = len(ligand.groups)
    decisions = set()
    for move in range(8):
        batch.multiple_steps(interval - 1)
        box_pre = batch.get_box(0)
        # the move comes after the Langevin update of this step; step alone to read it
        batch.multiple_steps(1)
        mv = batch.barostat_move(0)
        assert mv["triggered"] == 1.0
        x_post, box_post = batch.get_x_t(0), batch.get_box(0)
        V = float(np.prod(np.diag(box_pre)))
        w = obaro.acceptance_work(0.0, mv["delta_u"], V, mv["delta_volume"], nmol, pressure, 300.0)
        acc = obaro.accept(w, mv["r2"], 300.0)
        assert bool(mv["accepted"]) == acc, (move, w, mv)
        decisions.add(acc)
        ls = np.cbrt((V + mv["delta_volume"]) / V)
        assert mv["length_scale"] == ls
        if acc:
            np.testing.assert_array_equal(box_post, box_pre * np.where(np.eye(3) > 0, ls, 1.0))
            # u_after is the energy of the state the system keeps
            u = sum(bp.execute(x_post, box_post, 0.3)[2] for bp in bps)
        else:
            np.testing.assert_array_equal(box_post, box_pre)
            u = sum(bp.execute(x_post, box_post, 0.3)[2] for bp in bps)
        ref = mv["u_after"] if acc else mv["u_before"]
        assert abs(u - ref) <= 1e-9 * max(abs(ref), 1.0), (u, ref)
        if not acc:
            # the rejected image: oracle centroid scaling of the kept state
            x_img, box_img = obaro.scale_centroids(x_post, box_post, ligand.groups, ls)
            u_img = sum(bp.execute(x_img, box_img, 0.3)[2] for bp in bps)
            assert abs(u_img - mv["u_after"]) <= 1e-6 * abs(mv["u_after"])
    if pressure == 1.0:
        assert decisions == {True, False}


def test_no_move_between_intervals(ligand):
    ctx, keep = _context(ligand, 3, 5)[:2]
    batch = custom_ops.BatchedContexts([ctx], [0.0])
    batch.multiple_steps(3)
    assert batch.barostat_move(0)["triggered"] == 0.0


def test_ideal_gas_volume(ligand):
    """No nonbonded potential: the molecules are an ideal gas of flexible groups
    and <V> = N_mol kT / P, N_mol counting the ligand (the expectation of
    tests/test_barostat.py:262-365; the exact (N_mol + 1) kT / P of the
    isothermal-isobaric distribution differs by 1/N_mol = 0.05%). 100 bar; 300,
    600 and 1000 K; 2,000 moves each, the second half averaged; boxes start 2%
    above the target."""
    pressure = 100.0
    temps = [300.0, 600.0, 1000.0]
    interval = 5
    n_moves = 2000
    bonded = ligand.bound_potentials(include_nonbonded=False)
    nmol = len(ligand.groups)
    p_md = pressure * obaro.AVOGADRO * 1e-25
    expected = np.array([nmol * obaro.BOLTZ * T / p_md for T in temps])
    ctxs, keep = [], []
    for T, Vexp in zip(temps, expected):
        V0 = float(np.prod(np.diag(ligand.box)))
        ls = (1.02 * Vexp / V0) ** (1.0 / 3.0)
        x0, box0 = obaro.scale_centroids(ligand.x, ligand.box, ligand.groups, ls)
        bps = [p.bound_impl(F32) for p in bonded]
        c, k = _context(ligand, 2021, interval, pressure=pressure, temperature=T, bps=bps, box=box0, x=x0,
                        dt=0.0015)
        ctxs.append(c)
        keep.append(k)
    batch = custom_ops.BatchedContexts(ctxs, [0.0] * len(temps))
    vols = []
    per_call = 100  # moves per call: frames are stored with the boxes
    for _ in range(n_moves // per_call):
        _, _, boxes = batch.multiple_steps(per_call * interval, store_x_interval=interval)
        vols.append(np.prod(np.diagonal(boxes, axis1=2, axis2=3), axis=2))
    vols = np.concatenate(vols, axis=1)  # [S, n_moves]
    mean = vols[:, n_moves // 2 :].mean(axis=1)
    print("ideal gas <V> / expected:", (mean / expected).tolist(), flush=True)
    np.testing.assert_allclose(mean, expected, rtol=1e-2)


def test_partial_group_idxs(ligand):
    """Half of the molecules as barostat groups: the run is valid, volume moves
    are accepted, and the atoms outside every group keep their relative geometry
    through moves (only Langevin moves them, never the centroid scaling)."""
    groups = ligand.groups[len(ligand.groups) // 2 :]
    ctx, keep = _context(ligand, 2021, 3, groups=groups)[:2]
    _, _, boxes = ctx.multiple_steps(np.full(300, 1.0), store_x_interval=3)
    vols = np.prod(np.diagonal(boxes, axis1=1, axis2=2), axis=1)
    assert len(set(np.round(vols, 12))) > 1 and np.all(np.isfinite(ctx.get_x_t()))


def test_varying_pressure(ligand):
    """Pressure changes take effect (tests/test_barostat.py:196-259): from the
    same start, 1000 atm ends with a smaller box than 1 atm. Two identical
    contexts then continue from the same 1000 atm state, one switched to 1 atm
    with set_pressure (and set_interval, which restarts the call counter): it
    ends with the larger box. (The freshly built synthetic box first relaxes at
    any pressure, so volumes are compared with each other, not with the start.)"""
    atm = 1.01325
    hi, keep_hi = _context(ligand, 2021, 3, pressure=1000.0 * atm, dt=0.0015)[:2]
    hi2, keep_hi2 = _context(ligand, 2021, 3, pressure=1000.0 * atm, dt=0.0015)[:2]
    lo, keep_lo = _context(ligand, 2021, 3, pressure=1.0 * atm, dt=0.0015)[:2]
    for c in (hi, hi2, lo):
        c.multiple_steps(np.ones(1000))
    V_hi = float(np.prod(np.diag(hi.get_box())))
    V_lo = float(np.prod(np.diag(lo.get_box())))
    assert float(np.prod(np.diag(hi2.get_box()))) == V_hi  # identical copies
    print(f"after 1000 steps: V(1000 atm) {V_hi:.4f}, V(1 atm) {V_lo:.4f}", flush=True)
    assert V_lo - V_hi > 0.4
    keep_hi[2].set_pressure(1.0 * atm)
    keep_hi[2].set_interval(2)
    keep_hi2[2].set_interval(2)
    hi.multiple_steps(np.ones(2000))
    hi2.multiple_steps(np.ones(2000))
    V_sw = float(np.prod(np.diag(hi.get_box())))
    V_kept = float(np.prod(np.diag(hi2.get_box())))
    print(f"after 2000 more: switched to 1 atm {V_sw:.4f}, kept at 1000 atm {V_kept:.4f}", flush=True)
    assert V_sw - V_kept > 0.4


def test_move_stream_continues_across_engine_rebuilds(ligand):
    """30 steps in a Context, 30 in a batch with another system, 30 back in the
    Context: bitwise the 90 steps of an identical Context stepped alone. Every
    rebuild keeps the barostat's draws going (they are keyed on the absolute step)."""
    ref, keep_r = _context(ligand, 77, 3)[:2]
    dl_ref, _, _ = ref.multiple_steps(np.full(90, 0.4), store_du_dl_interval=1)
    a, keep_a = _context(ligand, 77, 3)[:2]
    other, keep_o = _context(ligand, 5, 3)[:2]
    d1, _, _ = a.multiple_steps(np.full(30, 0.4), store_du_dl_interval=1)
    batch = custom_ops.BatchedContexts([other, a], [0.9, 0.4])
    d2, _, _ = batch.multiple_steps(30, store_du_dl_interval=1)
    batch.write_back()
    del batch
    d3, _, _ = a.multiple_steps(np.full(30, 0.4), store_du_dl_interval=1)
    np.testing.assert_array_equal(np.concatenate([d1, d2[1], d3]), dl_ref)
    np.testing.assert_array_equal(a.get_x_t(), ref.get_x_t())
    np.testing.assert_array_equal(a.get_box(), ref.get_box())


def test_set_pressure_keeps_move_stream(ligand):
    """set_pressure rebuilds the context's engine (and resets the acceptance
    counters, barostat.cu:389-391); the next move draws new random numbers
    instead of replaying the first move's."""
    a, keep_a = _context(ligand, 123, 3)[:2]
    b1 = custom_ops.BatchedContexts([a], [0.2])
    b1.multiple_steps(3)
    m1 = b1.barostat_move(0)
    b1.multiple_steps(27)
    b1.write_back()
    del b1
    keep_a[2].set_pressure(1.0)
    b2 = custom_ops.BatchedContexts([a], [0.2])
    b2.multiple_steps(3)
    m2 = b2.barostat_move(0)
    assert m1["triggered"] == m2["triggered"] == 1.0
    assert m1["r2"] != m2["r2"]
