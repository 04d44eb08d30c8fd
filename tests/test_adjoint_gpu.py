"""Trajectory (pathwise) adjoint on the GPU: gradients of a loss on the du/dlambda
values of Langevin runs with respect to every forcefield parameter, by reverse
mode through the steps with one Hessian-vector product per step and frames
recomputed from checkpoints.

Checks:
  * against backpropagation through the same trajectory in Torch FP64 on the
    oracle energies (oracle.trajectory, itself pinned by finite differences in
    test_oracle.py), for a zero-temperature run (friction on, no noise, so the
    two trajectories are the same up to the force kernels' ~1e-7 error), with
    two lambda windows, a partial last checkpoint segment and per-frame loss
    weights: relative 1e-4 of each parameter block's largest entry;
  * the same at 300 K with noise, the oracle drawing the device's own Philox
    normals. (Finite differences of the GPU runs themselves are no check at
    these sizes: a pair crossing the unshifted cutoff between the perturbed
    runs makes the discrete loss jump, and the FP32 parameter values and
    approximations put ~1e-7 relative jitter on it; the oracle's derivative is
    pinned by finite differences on the CPU, on runs without such events);
  * bitwise repeatability, the state after the call, and the error paths.
"""

import dataclasses

import numpy as np
import pytest

from oracle import bonded as obonded
from oracle import integrator as ointeg
from oracle import nonbonded as onb
from oracle import philox as ophilox
from oracle import trajectory as otraj
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import custom_ops

pytestmark = pytest.mark.gpu
F32 = np.float32


def _small_system(seed=5):
    s = ts.ligand_in_water(side=2.0, n_waters=220, n_heavy=8, n_ligand=14, seed=seed)
    return dataclasses.replace(s, cutoff=0.8, nb_params=s.nb_params.astype(np.float32).astype(np.float64))


@pytest.fixture(scope="module")
def system(gpu):
    return _small_system()


def _batch(s, lams, dt, ca, cbs, ccs, seed=3, v0=None, bond_params=None, angle_params=None, torsion_params=None,
           nb_params=None):
    s2 = dataclasses.replace(
        s,
        bond_params=s.bond_params if bond_params is None else bond_params,
        angle_params=s.angle_params if angle_params is None else angle_params,
        torsion_params=s.torsion_params if torsion_params is None else torsion_params,
        nb_params=s.nb_params if nb_params is None else nb_params,
    )
    bps = [p.bound_impl(F32) for p in s2.bound_potentials()]
    W = len(lams)
    intg = custom_ops.LangevinIntegrator(dt, ca, cbs, ccs, seed)
    v0s = np.stack([v0 if v0 is not None else np.zeros((s.N, 3))] * W)
    return custom_ops.BatchedContexts.from_windows(s.x, v0s, s.box, intg, [seed + w for w in range(W)], bps, lams)


def _oracle_energy(s, margin=0.5):
    """The oracle energy over a pair list fixed at the start, with a margin the
    fastest atoms do not cross in these runs (0.15 nm was crossed at ~30 steps)."""
    pairs = onb.candidate_pairs(s.x, s.box, s.cutoff + margin)
    excluded = {tuple(p) for p in np.sort(np.asarray(s.exclusion_idxs), axis=1).tolist()}
    keep = np.array([tuple(p) not in excluded for p in pairs.tolist()], dtype=bool)
    none = np.zeros((0, 2), int)
    nb_pairs = onb.pair_energy_fn(s.x, s.box, none, np.zeros((0, 2)), s.beta, s.cutoff, s.lambda_plane_idxs,
                                  s.lambda_offset_idxs, pairs=pairs[keep])
    nb_rest = onb.pair_list_energy_fn(s.exclusion_idxs, 1.0 - np.asarray(s.scales), s.box, s.beta, s.cutoff,
                                      s.lambda_plane_idxs, s.lambda_offset_idxs)
    fb = obonded.bonded_energy_fn("bond", s.bond_idxs)
    fa = obonded.bonded_energy_fn("angle", s.angle_idxs)
    ft = obonded.bonded_energy_fn("torsion", s.torsion_idxs)

    def energy(x, ps, lt):  # ps = [nonbonded, bond, angle, torsion]: the engine's flat order
        return nb_pairs(x, ps[0], lt) + nb_rest(x, ps[0], lt) + fb(x, ps[1], lt) + fa(x, ps[2], lt) + ft(x, ps[3], lt)

    return energy


def _blocks(s):
    return [s.nb_params, s.bond_params, s.angle_params, s.torsion_params]


def test_adjoint_matches_fp64_backprop_through_the_trajectory(system):
    s = system
    dt, friction = 0.0015, 1.0
    ca, cb, _ = ointeg.langevin_coefficients(300.0, dt, friction, s.masses)
    cbs = -cb
    ccs = np.zeros(s.N)  # no noise: the oracle follows the same trajectory
    lams = [0.3, 0.7]
    n_steps, interval, ck = 24, 3, 9  # 24 = 9 + 9 + 6: a partial last segment
    F = n_steps // interval
    rng = np.random.default_rng(7)
    v0 = rng.normal(size=(s.N, 3)) * np.sqrt(2.49 / s.masses)[:, None]
    batch = _batch(s, lams, dt, ca, cbs, ccs, v0=v0)
    dl = batch.adjoint_forward(n_steps, interval, ck)
    assert dl.shape == (2, F)
    dl_bar = rng.normal(size=(2, F))
    grads, xb, vb = batch.adjoint_backward(dl_bar, want_initial_state=True)
    energy = _oracle_energy(s)
    N = s.N
    for w, lam in enumerate(lams):
        rdl, rgrads, rxb, rvb = otraj.trajectory_loss_grad(energy, s.x, v0, _blocks(s), lam, n_steps, interval, dt,
                                                           ca, cbs, ccs, dl_bar[w])
        err_dl = np.abs(dl[w] - rdl).max() / max(np.abs(rdl).max(), 1.0)
        print(f"[window {w}] du_dl max rel err {err_dl:.2e}", flush=True)
        assert err_dl <= 1e-5
        o = 0
        for name, ref in zip(["nonbonded", "bond", "angle", "torsion"], rgrads):
            got = grads[w][o:o + ref.size].reshape(ref.shape)
            o += ref.size
            err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-12)
            print(f"[window {w}] d loss / d {name}: max err / max|ref| = {err:.2e}", flush=True)
            assert err <= 1e-4
        assert o == grads[w].size
        for name, got, ref in [("x0", xb[w * N:(w + 1) * N], rxb), ("v0", vb[w * N:(w + 1) * N], rvb)]:
            err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-12)
            print(f"[window {w}] d loss / d {name}: max err / max|ref| = {err:.2e}", flush=True)
            assert err <= 1e-4


def test_adjoint_with_noise_matches_fp64_backprop(system):
    """300 K, friction 1/ps: the oracle follows the GPU trajectory with the
    device's own Langevin noise (oracle.philox, tied to the device header in
    test_host_math.py), so the comparison of test 1 holds with noise on."""
    s = system
    dt = 0.0015
    ca, cb, cc = ointeg.langevin_coefficients(300.0, dt, 1.0, s.masses)
    cbs = -cb
    n_steps, interval, ck, seed = 40, 4, 16, 3  # 40 = 16 + 16 + 8
    F = n_steps // interval
    rng = np.random.default_rng(9)
    dl_bar = rng.normal(size=(1, F))
    b = _batch(s, [0.5], dt, ca, cbs, cc, seed=seed)  # window 0 draws with seed `seed`
    dl = b.adjoint_forward(n_steps, interval, ck)
    (g,), xb, vb = b.adjoint_backward(dl_bar, want_initial_state=True)
    noise = np.stack([ophilox.langevin_noise(seed, j, s.N) for j in range(n_steps)])
    rdl, rgrads, rxb, rvb = otraj.trajectory_loss_grad(_oracle_energy(s), s.x, np.zeros((s.N, 3)), _blocks(s), 0.5,
                                                       n_steps, interval, dt, ca, cbs, cc, dl_bar[0], noise)
    err_dl = np.abs(dl[0] - rdl).max() / max(np.abs(rdl).max(), 1.0)
    print(f"[300 K] du_dl max rel err {err_dl:.2e}", flush=True)
    assert err_dl <= 1e-5
    o = 0
    for name, ref in zip(["nonbonded", "bond", "angle", "torsion"], rgrads):
        got = g[o:o + ref.size].reshape(ref.shape)
        o += ref.size
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-12)
        print(f"[300 K] d loss / d {name}: max err / max|ref| = {err:.2e}", flush=True)
        assert err <= 1e-4
    for name, got, ref in [("x0", xb, rxb), ("v0", vb, rvb)]:
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-12)
        print(f"[300 K] d loss / d {name}: max err / max|ref| = {err:.2e}", flush=True)
        assert err <= 1e-4


def test_adjoint_is_bitwise_repeatable_and_leaves_the_end_state(system):
    s = system
    dt = 0.0015
    ca, cb, cc = ointeg.langevin_coefficients(300.0, dt, 1.0, s.masses)
    b = _batch(s, [0.2, 0.9], dt, ca, -cb, cc)
    b.adjoint_forward(30, 5, 10)
    x_end = [b.get_x_t(w).copy() for w in range(2)]
    v_end = [b.get_v_t(w).copy() for w in range(2)]
    dl_bar = np.ones((2, 6))
    g1 = b.adjoint_backward(dl_bar)
    g2 = b.adjoint_backward(dl_bar)
    for a, c in zip(g1, g2):
        np.testing.assert_array_equal(a, c)
    for w in range(2):
        np.testing.assert_array_equal(b.get_x_t(w), x_end[w])
        np.testing.assert_array_equal(b.get_v_t(w), v_end[w])
    # continuing from there equals a batch that ran 30 + 10 steps straight
    ref = _batch(s, [0.2, 0.9], dt, ca, -cb, cc)
    ref.multiple_steps(40)
    b.multiple_steps(10)
    for w in range(2):
        np.testing.assert_array_equal(b.get_x_t(w), ref.get_x_t(w))


def test_adjoint_error_paths(system):
    s = system
    dt = 0.0015
    ca, cb, cc = ointeg.langevin_coefficients(300.0, dt, 1.0, s.masses)
    b = _batch(s, [0.5], dt, ca, -cb, cc)
    with pytest.raises(RuntimeError, match="adjoint_forward first"):
        b.adjoint_backward(np.ones((1, 1)))
    with pytest.raises(RuntimeError, match="multiple"):
        b.adjoint_forward(10, 3, 4)


@pytest.mark.parametrize("recompute", [True, False])
def test_ti_loss_adjoint_method_matches_fp64_backprop(system, recompute):
    """The torch.autograd.Function over lambda windows with method="adjoint":
    the gradient of L = (sum_w c_w <du/dlambda>_w - target)^2 (equilibration
    steps included in the differentiated run) in the nonbonded parameters and
    torsion k, against FP64 backpropagation through each window's trajectory
    with its own noise, summed over the windows."""
    import torch

    from timemachine_amd.fe import window_gradients as wg

    s = system
    lams = [0.25, 0.75]
    n_frames, interval, equil = 4, 3, 6
    windows = wg.LambdaWindows(s, lams, n_frames, interval, equil_steps=equil, dt=0.0015, seed=11)
    nb = torch.tensor(s.nb_params, dtype=torch.float64, requires_grad=True)
    tk = torch.tensor(s.torsion_params[:, 0], dtype=torch.float64, requires_grad=True)
    target = -3.0
    loss = wg.ti_loss(nb, tk, windows, target, recompute=recompute, method="adjoint")
    g_nb, g_tk = torch.autograd.grad(loss, (nb, tk))
    # oracle
    ca, cb, cc = ointeg.langevin_coefficients(windows.temperature, windows.dt, windows.friction, s.masses)
    energy = _oracle_energy(s)
    c = wg.trapezoid_weights(lams)
    n_steps = equil + n_frames * interval
    means, ref_grads = [], []
    for w, lam in enumerate(lams):
        seed = windows.seed + 7919 * w
        noise = np.stack([ophilox.langevin_noise(seed, j, s.N) for j in range(n_steps)])
        dls = otraj.trajectory_du_dl(energy, windows.x0, windows.v0[w], _blocks(s), lam, n_steps, interval,
                                     windows.dt, ca, -cb, cc, noise)
        means.append(dls[equil // interval:].mean())
    lbar = 2.0 * (float(np.dot(c, means)) - target)
    for w, lam in enumerate(lams):
        seed = windows.seed + 7919 * w
        noise = np.stack([ophilox.langevin_noise(seed, j, s.N) for j in range(n_steps)])
        dl_bar = np.zeros(n_steps // interval)
        dl_bar[equil // interval:] = lbar * c[w] / n_frames
        _, grads, _, _ = otraj.trajectory_loss_grad(energy, windows.x0, windows.v0[w], _blocks(s), lam, n_steps,
                                                    interval, windows.dt, ca, -cb, cc, dl_bar, noise)
        ref_grads.append(grads)
    r_nb = sum(g[0] for g in ref_grads)
    r_tk = sum(g[3][:, 0] for g in ref_grads)
    lv = float(loss.detach())
    assert abs(lv - (float(np.dot(c, means)) - target) ** 2) <= 1e-5 * max(lv, 1.0)
    for name, got, ref in [("nonbonded", g_nb.numpy(), r_nb), ("torsion k", g_tk.numpy(), r_tk)]:
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-12)
        print(f"[ti_loss adjoint recompute={recompute}] d loss / d {name}: max err / max|ref| = {err:.2e}",
              flush=True)
        assert err <= 1e-4


def test_truncated_adjoint_matches_truncated_backprop(system):
    """horizon = 6: the reverse sweep's state cotangents cleared every 6 steps
    equals FP64 backpropagation with the state's history cut at the same
    steps; horizon >= n_steps is the full sweep, bitwise."""
    s = system
    dt = 0.0015
    ca, cb, _ = ointeg.langevin_coefficients(300.0, dt, 1.0, s.masses)
    ccs = np.zeros(s.N)
    n_steps, interval, ck, H = 24, 3, 9, 6
    rng = np.random.default_rng(13)
    v0 = rng.normal(size=(s.N, 3)) * np.sqrt(2.49 / s.masses)[:, None]
    b = _batch(s, [0.4], dt, ca, -cb, ccs, v0=v0)
    b.adjoint_forward(n_steps, interval, ck)
    dl_bar = rng.normal(size=(1, n_steps // interval))
    (g_t,) = b.adjoint_backward(dl_bar, horizon=H)
    (g_full,) = b.adjoint_backward(dl_bar)
    (g_long,) = b.adjoint_backward(dl_bar, horizon=n_steps)
    np.testing.assert_array_equal(g_long, g_full)
    _, rgrads, _, _ = otraj.trajectory_loss_grad(_oracle_energy(s), s.x, v0, _blocks(s), 0.4, n_steps, interval, dt,
                                                 ca, -cb, ccs, dl_bar[0], detach_every=H)
    o = 0
    for name, ref in zip(["nonbonded", "bond", "angle", "torsion"], rgrads):
        got = g_t[o:o + ref.size].reshape(ref.shape)
        o += ref.size
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-12)
        print(f"[horizon {H}] d loss / d {name}: max err / max|ref| = {err:.2e}", flush=True)
        assert err <= 1e-4
    assert np.abs(g_t - g_full).max() > 0.0
