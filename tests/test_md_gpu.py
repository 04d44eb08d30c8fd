"""Integrator, barostat, Context and batched stepping on the GPU.

* one BAOAB step reproduces the FP64 update formula bit for bit from the
  step's own fixed-point forces (integrator.cu:37-69);
* with friction 0 the noise vanishes and the trajectory matches the oracle's
  leapfrog driven by oracle forces to the force tolerance;
* a system stepped inside a batch is bitwise identical to the same system
  stepped alone (barostat on), and run-to-run identical.
"""

import numpy as np
import pytest

from oracle import integrator as oint
from oracle import nonbonded as onb
from timemachine_amd import _lib
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import LangevinIntegrator, MonteCarloBarostat, custom_ops

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def ligand(gpu):
    return ts.ligand_in_water()


def _context(sys, seed=2022, friction=1.0, dt=0.0025, barostat_interval=0, temperature=300.0):
    bps = [p.bound_impl(F32) for p in sys.bound_potentials()]
    intg = LangevinIntegrator(temperature, dt, friction, sys.masses, seed).impl()
    v0 = ts.maxwell_boltzmann(sys.masses, temperature, seed)
    baro = None
    if barostat_interval:
        baro = MonteCarloBarostat(sys.N, 1.0, temperature, sys.groups, barostat_interval, seed + 1).impl(bps)
    ctx = custom_ops.Context(sys.x, v0, sys.box, intg, bps, barostat=baro)
    return ctx, (bps, intg, baro), v0


def test_single_step_matches_update_formula(ligand):
    ctx, keep, v0 = _context(ligand, friction=0.0)
    ctx.step(0.5)
    f = ctx._get_du_dx_t_minus_1()
    intg = LangevinIntegrator(300.0, 0.0025, 0.0, ligand.masses, 2022)
    v_mid = v0 + intg.cbs[:, None] * f
    v_new = intg.ca * v_mid + 0.0
    x_new = ligand.x + 0.5 * 0.0025 * (v_mid + v_new)
    np.testing.assert_array_equal(ctx.get_v_t(), v_new)
    np.testing.assert_array_equal(ctx.get_x_t(), x_new)


def test_forces_in_context_match_oracle(ligand):
    ctx, keep, _ = _context(ligand)
    ctx.step(0.5)
    f = ctx._get_du_dx_t_minus_1()
    from oracle import bonded as ob

    ref = onb.nonbonded_pairs(ligand.x, ligand.nb_params, ligand.box, 0.5, ligand.exclusion_idxs, ligand.scales,
                              2.0, 1.2, ligand.lambda_plane_idxs, ligand.lambda_offset_idxs)[0]
    ref = ref + ob.bonded("bond", ligand.x, ligand.bond_params, 0.5, ligand.bond_idxs)[0]
    ref = ref + ob.bonded("angle", ligand.x, ligand.angle_params, 0.5, ligand.angle_idxs)[0]
    ref = ref + ob.bonded("torsion", ligand.x, ligand.torsion_params, 0.5, ligand.torsion_idxs)[0]
    err = onb.force_rel_err(f, ref)
    print(f"total force rel err {err.max():.3e}")
    assert err.max() <= 1e-5


def test_friction_zero_trajectory_tracks_oracle(ligand):
    """20 NVE steps: GPU trajectory vs oracle leapfrog with oracle forces."""
    ctx, keep, v0 = _context(ligand, friction=0.0)
    n = 20
    _, xs, _ = ctx.multiple_steps(np.full(n, 0.2), store_du_dl_interval=0, store_x_interval=1)
    from oracle import bonded as ob

    intg = LangevinIntegrator(300.0, 0.0025, 0.0, ligand.masses, 2022)
    x, v = ligand.x.copy(), v0.copy()
    for k in range(n):
        f = onb.nonbonded_pairs(x, ligand.nb_params, ligand.box, 0.2, ligand.exclusion_idxs, ligand.scales, 2.0, 1.2,
                                ligand.lambda_plane_idxs, ligand.lambda_offset_idxs)[0]
        f += ob.bonded("bond", x, ligand.bond_params, 0.2, ligand.bond_idxs)[0]
        f += ob.bonded("angle", x, ligand.angle_params, 0.2, ligand.angle_idxs)[0]
        f += ob.bonded("torsion", x, ligand.torsion_params, 0.2, ligand.torsion_idxs)[0]
        x, v = oint.baoab_step(x, v, f, 0.0025, 1.0, intg.cbs, np.zeros(ligand.N), np.zeros_like(x))
        assert np.abs(xs[k] - x).max() < 1e-6 * (k + 1)


def test_context_store_intervals(ligand):
    ctx, keep, _ = _context(ligand)
    du_dl, xs, boxes = ctx.multiple_steps(np.full(12, 0.5), store_du_dl_interval=3, store_x_interval=4)
    assert du_dl.shape == (4,) and xs.shape == (3, ligand.N, 3) and boxes.shape == (3, 3, 3)
    assert np.all(np.isfinite(du_dl)) and np.all(du_dl != 0)
    np.testing.assert_array_equal(xs[-1], ctx.get_x_t())
    du_dl0, xs0, _ = ctx.multiple_steps(np.full(5, 0.5))  # intervals 0 -> only the last step
    assert du_dl0.shape == (1,) and xs0.shape == (1, ligand.N, 3)


def test_run_to_run_and_batch_bitwise(ligand):
    """Each system of a batch (barostat on) equals itself stepped alone, bit for bit."""
    lams = [0.0, 0.45, 1.0]
    alone = []
    for s, lam in enumerate(lams):
        ctx, keep, _ = _context(ligand, seed=100 + s, barostat_interval=5)
        du_dl, xs, boxes = ctx.multiple_steps(np.full(30, lam), store_du_dl_interval=1, store_x_interval=10)
        alone.append((du_dl, xs, boxes, ctx.get_x_t(), ctx.get_v_t(), ctx.get_box()))
    ctxs = []
    keeps = []
    for s, lam in enumerate(lams):
        ctx, keep, _ = _context(ligand, seed=100 + s, barostat_interval=5)
        ctxs.append(ctx)
        keeps.append(keep)
    batch = custom_ops.BatchedContexts(ctxs, lams)
    du_dl, frames, boxes = batch.multiple_steps(30, store_du_dl_interval=1, store_x_interval=10)
    for s in range(len(lams)):
        np.testing.assert_array_equal(du_dl[s], alone[s][0])
        np.testing.assert_array_equal(frames[s], alone[s][1])
        np.testing.assert_array_equal(boxes[s], alone[s][2])
        np.testing.assert_array_equal(batch.get_x_t(s), alone[s][3])
        np.testing.assert_array_equal(batch.get_v_t(s), alone[s][4])
        np.testing.assert_array_equal(batch.get_box(s), alone[s][5])


@pytest.mark.parametrize("use_graph", [True, False])
def test_deferred_barostat_moves_bitwise(ligand, use_graph):
    """Inside one multiple_steps call a barostat move is evaluated together with
    the next step's forces; one step per call never defers. Both give the same
    trajectory, du/dlambda series and boxes, bit for bit."""
    lams = [0.0, 0.6]

    def batch():
        made = [_context(ligand, seed=300 + s, barostat_interval=3) for s in range(len(lams))]
        return custom_ops.BatchedContexts([m[0] for m in made], lams), made

    b1, keep1 = batch()
    dl1, _, _ = b1.multiple_steps(31, store_du_dl_interval=1, use_graph=use_graph)
    b2, keep2 = batch()
    dl2 = np.concatenate([b2.multiple_steps(1, store_du_dl_interval=1, use_graph=use_graph)[0] for _ in range(31)],
                         axis=1)
    np.testing.assert_array_equal(dl1, dl2)
    vols = set()
    for s in range(len(lams)):
        np.testing.assert_array_equal(b1.get_x_t(s), b2.get_x_t(s))
        np.testing.assert_array_equal(b1.get_v_t(s), b2.get_v_t(s))
        np.testing.assert_array_equal(b1.get_box(s), b2.get_box(s))
        vols.add(float(np.prod(np.diag(b1.get_box(s)))))
    assert all(abs(v - 4.0**3) > 1e-9 for v in vols)  # moves were accepted


def test_deferred_barostat_moves_with_lambda_schedule_bitwise(ligand):
    """Context.multiple_steps with a lambda schedule that holds for a few steps and
    then jumps: moves are deferred only where lambda does not change. One call and
    one step per call agree bit for bit."""
    sched = np.repeat([0.2, 0.5, 0.8, 0.35, 0.9], 7)[:31]
    c1, keep1, _ = _context(ligand, seed=41, barostat_interval=3)
    dl1, _, _ = c1.multiple_steps(sched, store_du_dl_interval=1)
    c2, keep2, _ = _context(ligand, seed=41, barostat_interval=3)
    dl2 = np.concatenate([c2.multiple_steps(sched[j : j + 1], store_du_dl_interval=1)[0] for j in range(len(sched))])
    np.testing.assert_array_equal(dl1, dl2)
    np.testing.assert_array_equal(c1.get_x_t(), c2.get_x_t())
    np.testing.assert_array_equal(c1.get_v_t(), c2.get_v_t())
    np.testing.assert_array_equal(c1.get_box(), c2.get_box())


def test_graph_and_eager_paths_agree(ligand):
    ctxs = [_context(ligand, seed=7, barostat_interval=5)[0] for _ in range(2)]
    b1 = custom_ops.BatchedContexts([ctxs[0]], [0.3])
    b2 = custom_ops.BatchedContexts([ctxs[1]], [0.3])
    r1 = b1.multiple_steps(25, store_du_dl_interval=1, use_graph=True)
    r2 = b2.multiple_steps(25, store_du_dl_interval=1, use_graph=False)
    np.testing.assert_array_equal(r1[0], r2[0])
    np.testing.assert_array_equal(b1.get_x_t(0), b2.get_x_t(0))


def test_graphs_kept_across_runs_equal_eager(ligand):
    """An engine keeps the graph executables of runs without frames or a lambda
    schedule and replays them in later calls. Runs of several lengths (the
    du/dl buffer grows past its first size, which drops the kept graphs), a
    run without du/dl, a change of tile grid and occupancy build in between
    (new graphs), all against the eager path: du/dl series and final states
    bit for bit."""
    lams = [0.25, 0.75]

    def batch():
        made = [_context(ligand, seed=500 + s, barostat_interval=5) for s in range(len(lams))]
        return custom_ops.BatchedContexts([m[0] for m in made], lams), made

    bg, keep_g = batch()
    be, keep_e = batch()
    plan = [(17, 1), (30, 1), (45, 5), (20, 0), (130, 1), (30, 1), (30, 1)]
    for k, (n, interval) in enumerate(plan):
        if k == 5:
            for b in (bg, be):
                b.set_tile_grid(4096)
                b.set_tile_waves(3)
        print(f"run {k}: {n} steps, du/dl every {interval}", flush=True)
        dg, _, _ = bg.multiple_steps(n, store_du_dl_interval=interval, use_graph=True)
        de, _, _ = be.multiple_steps(n, store_du_dl_interval=interval, use_graph=False)
        np.testing.assert_array_equal(dg, de)
        assert dg.shape == (len(lams), n // interval if interval else 0)
    for s in range(len(lams)):
        np.testing.assert_array_equal(bg.get_x_t(s), be.get_x_t(s))
        np.testing.assert_array_equal(bg.get_v_t(s), be.get_v_t(s))
        np.testing.assert_array_equal(bg.get_box(s), be.get_box(s))


def test_barostat_moves_volume_and_validates(ligand):
    ctx, keep, _ = _context(ligand, barostat_interval=3)
    v0 = np.prod(np.diag(ctx.get_box()))
    _, _, boxes = ctx.multiple_steps(np.full(60, 0.0), store_x_interval=3)
    vols = np.prod(np.diagonal(boxes, axis1=1, axis2=2), axis=1)
    assert len(set(np.round(vols, 12))) > 1  # some moves accepted
    # the freshly built box relaxes (expands by up to ~8%) within the first few hundred steps
    assert np.all(np.abs(vols / v0 - 1) < 0.10)
    bps = keep[0]
    with pytest.raises(RuntimeError, match="Barostat interval must be greater than 0"):
        MonteCarloBarostat(ligand.N, 1.0, 300.0, ligand.groups, 0, 1).impl(bps)
    b = keep[2]
    assert b.get_interval() == 3
    b.set_interval(7)
    assert b.get_interval() == 7
    ctx.multiple_steps(np.full(7, 0.0))


def test_context_rejects_mismatched_velocities(ligand):
    bps = [p.bound_impl(F32) for p in ligand.bound_potentials()]
    intg = LangevinIntegrator(300.0, 0.0025, 1.0, ligand.masses, 1).impl()
    with pytest.raises(RuntimeError, match="v0 N != x0 N"):
        custom_ops.Context(ligand.x, np.zeros((3, 3)), ligand.box, intg, bps)


def test_engine_forces_equal_potential_sum_bitwise(ligand):
    """The engine gathers bonded forces per atom and evaluates the nonbonded
    terms in one batch; its fixed-point du_dx equals the sum of the separate
    Potential executions (term-by-term bonded kernels) bit for bit."""
    ctx, keep, _ = _context(ligand)
    ctx.step(0.5)
    raw = np.zeros(ligand.N * 3, dtype=np.uint64)
    _lib.call("tm_context_get_du_dx_t_minus_1", ctx.handle, _lib.pu64(raw))
    total = np.zeros(ligand.N * 3, dtype=np.uint64)
    for bp in ligand.bound_potentials():
        pot = bp.unbound_impl(np.float32)
        total += pot.execute_fixed_raw(ligand.x, bp.params, ligand.box, 0.5)[0]  # modulo 2^64
    np.testing.assert_array_equal(raw, total)


def test_empty_schedule_raises(ligand):
    """An empty lambda schedule resolves the store intervals to 0, which the
    reference rejects (context.cu:60-65)."""
    ctx, keep, _ = _context(ligand)
    with pytest.raises(RuntimeError, match="store_du_dl_interval <= 0"):
        ctx.multiple_steps(np.zeros(0))
    with pytest.raises(RuntimeError, match="store_u_interval <= 0"):
        ctx.multiple_steps_U(0.0, 0, [0.0, 1.0], 0, 0)


def test_engine_rejects_pair_lists_with_different_planes(ligand):
    """Pair lists of one system share one per-atom plane / offset array in the
    engine: a second list with other lambda offsets is refused rather than
    evaluated with the first list's w coordinates."""
    from timemachine_amd.lib import potentials

    pl1 = potentials.NonbondedPairListNegated(ligand.exclusion_idxs, ligand.scales, ligand.lambda_plane_idxs,
                                              ligand.lambda_offset_idxs, 2.0, 1.2).bind(ligand.nb_params)
    other_offsets = np.zeros_like(ligand.lambda_offset_idxs)
    pl2 = potentials.NonbondedPairList(ligand.exclusion_idxs[:5], ligand.scales[:5], ligand.lambda_plane_idxs,
                                       other_offsets, 2.0, 1.2).bind(ligand.nb_params)
    bps = [pl1.bound_impl(F32), pl2.bound_impl(F32)]
    intg = LangevinIntegrator(300.0, 0.0025, 1.0, ligand.masses, 1).impl()
    with pytest.raises(RuntimeError, match="must share parameters, lambda plane and offset"):
        custom_ops.Context(ligand.x, np.zeros_like(ligand.x), ligand.box, intg, bps)


def test_build_pieces_do_not_change_bits(ligand):
    """The neighborlist build split over 1, 2, 4 or 8 waves per row block (the
    pieces' leftovers merged into tiles) gives the same tile count and bitwise
    the same trajectories, barostat on."""
    lams = [0.2, 0.6, 0.9]
    ref = None
    for pieces in (1, 2, 4, 8):
        made = [_context(ligand, seed=700 + s, barostat_interval=5) for s in range(len(lams))]
        b = custom_ops.BatchedContexts([m[0] for m in made], lams)
        b.set_build_pieces(pieces)
        du_dl, _, _ = b.multiple_steps(40, store_du_dl_interval=1)
        out = (du_dl, [b.get_x_t(s) for s in range(len(lams))], b.total_tiles())
        print(f"pieces {pieces}: tiles {out[2]}", flush=True)
        if ref is None:
            ref = out
        else:
            np.testing.assert_array_equal(out[0], ref[0])
            for a, c in zip(out[1], ref[1]):
                np.testing.assert_array_equal(a, c)
            assert out[2] == ref[2]
    with pytest.raises(RuntimeError, match="pieces"):
        b.set_build_pieces(3)
