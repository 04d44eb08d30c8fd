"""Hessian-vector products (the du_dx adjoint) on the GPU against the FP64 oracle.

execute_hvp returns H (vx, vp, vl) for H the Hessian of U in (coords,
params, lambda): the tangents of du_dx, du_dp and du_dl. The oracle takes the
same product by Torch FP64 double backward (pinned in test_oracle.py by
central differences of its gradient, a closed-form bond Hessian and
symmetry). The kernels evaluate exact FP64 derivatives of the force kernels'
energy at the FP32 values of the nonbonded parameters (the test system's
parameters are rounded to FP32 here, so both sides see the same values), so
the tolerance is set by the 2^-36 fixed-point accumulation:
per-atom |dh| / max(|h|, 1) <= 1e-8, du_dp tangents per column <= 1e-8 of
the column's largest, du_dl tangent <= 1e-8 relative. Results are bitwise
run to run (fixed point, order independent).
"""

import dataclasses

import numpy as np
import pytest

from oracle import bonded as obonded
from oracle import nonbonded as onb
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import potentials

pytestmark = pytest.mark.gpu
F32 = np.float32
TOL = 1e-8


@pytest.fixture(scope="module")
def ligand(gpu):
    s = ts.ligand_in_water(side=3.0, n_waters=800, n_heavy=12, n_ligand=22, seed=3)
    return dataclasses.replace(s, nb_params=s.nb_params.astype(np.float32).astype(np.float64))


def _tangents(sys, rng, ligand_scale=1.0):
    """Random tangents. ligand_scale shrinks them on the ligand: the LJ
    Hessian of a bonded (excluded) ligand pair at 0.1 nm is ~1e10, beyond the
    2^63 / 2^36 ~ 1.3e8 range of the fixed-point accumulators. In the full
    Nonbonded potential the all-pairs and exclusion terms of such a pair are
    bitwise negatives and cancel (test_nonbonded_hvp); the parts alone need
    tangents that keep them in range."""
    vx = rng.normal(size=(sys.N, 3))
    vp = rng.normal(size=sys.nb_params.shape) * np.array([1.0, 0.01, 0.1])
    # LJ is off where eps == 0 (gradients included): keep those eps at 0
    vp[sys.nb_params[:, 2] == 0.0, 2] = 0.0
    lig = np.asarray(sys.ligand_idxs)
    vx[lig] *= ligand_scale
    vp[lig] *= ligand_scale
    return vx, vp


def _check(out, ref, label):
    hx, hp, hl = out
    rx, rp, rl = ref
    ex = onb.force_rel_err(hx, rx).max()
    print(f"[{label}] hx max rel err {ex:.3e}  hl {hl:.9g} ref {rl:.9g}", flush=True)
    assert ex <= TOL
    rp = np.asarray(rp).reshape(np.asarray(hp).shape)
    for c in range(rp.shape[1]):
        scale = max(np.abs(rp[:, c]).max(), 1e-12)
        err = np.abs(hp[:, c] - rp[:, c]).max() / scale
        print(f"[{label}] hp column {c} max err / max|ref| = {err:.3e}", flush=True)
        assert err <= TOL
    assert abs(hl - rl) <= TOL * max(abs(rl), 1.0)


@pytest.mark.parametrize("lam", [0.0, 0.35, 1.0])
def test_all_pairs_hvp(ligand, lam):
    """NonbondedAllPairs alone (no exclusions): every pair within the cutoff."""
    rng = np.random.default_rng(11)
    s = ligand
    pot = potentials.NonbondedAllPairs(s.lambda_plane_idxs, s.lambda_offset_idxs, s.beta, s.cutoff).unbound_impl(F32)
    vx, vp = _tangents(s, rng, ligand_scale=1e-4)
    out = pot.execute_hvp(s.x, s.nb_params, s.box, lam, vx, vp, 0.8)
    fn = onb.pair_energy_fn(s.x, s.box, np.zeros((0, 2), int), np.zeros((0, 2)), s.beta, s.cutoff,
                            s.lambda_plane_idxs, s.lambda_offset_idxs)
    _check(out, onb.hvp(fn, s.x, s.nb_params, lam, vx, vp, 0.8), f"all pairs lam={lam}")


def test_exclusions_hvp(ligand):
    """NonbondedPairList_negated over the exclusions with their scales."""
    rng = np.random.default_rng(12)
    s = ligand
    pot = potentials.NonbondedPairListNegated(s.exclusion_idxs, s.scales, s.lambda_plane_idxs, s.lambda_offset_idxs,
                                              s.beta, s.cutoff).unbound_impl(F32)
    vx, vp = _tangents(s, rng, ligand_scale=1e-4)
    out = pot.execute_hvp(s.x, s.nb_params, s.box, 0.35, vx, vp, -0.4)
    fn = onb.pair_list_energy_fn(s.exclusion_idxs, s.scales, s.box, s.beta, s.cutoff, s.lambda_plane_idxs,
                                 s.lambda_offset_idxs, negated=True)
    _check(out, onb.hvp(fn, s.x, s.nb_params, 0.35, vx, vp, -0.4), "exclusions")


def test_nonbonded_hvp(ligand):
    """The full Nonbonded potential (all pairs minus scaled exclusions)."""
    rng = np.random.default_rng(13)
    s = ligand
    pot = s.nonbonded().unbound_impl(F32)
    vx, vp = _tangents(s, rng)
    out = pot.execute_hvp(s.x, s.nb_params, s.box, 0.5, vx, vp, 1.0)
    # the same energy without cancellation: the pairs within the cutoff that are
    # not excluded, plus each exclusion at (1 - scale); the FP64 difference of
    # the all-pairs and exclusion terms of a bonded pair (~1e10) would leave
    # ~1e-6 of rounding in the oracle, where the fixed-point sums cancel exactly
    pairs = onb.candidate_pairs(s.x, s.box, s.cutoff)
    excluded = {tuple(p) for p in np.sort(np.asarray(s.exclusion_idxs), axis=1).tolist()}
    keep = np.array([tuple(p) not in excluded for p in pairs.tolist()], dtype=bool)
    none = np.zeros((0, 2), int)
    fn_pairs = onb.pair_energy_fn(s.x, s.box, none, np.zeros((0, 2)), s.beta, s.cutoff, s.lambda_plane_idxs,
                                  s.lambda_offset_idxs, pairs=pairs[keep])
    fn_rest = onb.pair_list_energy_fn(s.exclusion_idxs, 1.0 - np.asarray(s.scales), s.box, s.beta, s.cutoff,
                                      s.lambda_plane_idxs, s.lambda_offset_idxs)

    def fn(xt, pt, lt):
        return fn_pairs(xt, pt, lt) + fn_rest(xt, pt, lt)

    _check(out, onb.hvp(fn, s.x, s.nb_params, 0.5, vx, vp, 1.0), "nonbonded")


def test_interaction_group_hvp(ligand):
    """NonbondedInteractionGroup: ligand rows against every other atom."""
    rng = np.random.default_rng(14)
    s = ligand
    rows = np.asarray(s.ligand_idxs, dtype=np.int32)
    pot = potentials.NonbondedInteractionGroup(rows, s.lambda_plane_idxs, s.lambda_offset_idxs, s.beta,
                                               s.cutoff).unbound_impl(F32)
    vx, vp = _tangents(s, rng)
    out = pot.execute_hvp(s.x, s.nb_params, s.box, 0.2, vx, vp, 0.5)
    is_row = np.zeros(s.N, bool)
    is_row[rows] = True
    pairs = onb.candidate_pairs(s.x, s.box, s.cutoff)
    pairs = pairs[is_row[pairs[:, 0]] != is_row[pairs[:, 1]]]
    fn = onb.pair_energy_fn(s.x, s.box, np.zeros((0, 2), int), np.zeros((0, 2)), s.beta, s.cutoff,
                            s.lambda_plane_idxs, s.lambda_offset_idxs, pairs=pairs)
    _check(out, onb.hvp(fn, s.x, s.nb_params, 0.2, vx, vp, 0.5), "interaction group")


@pytest.mark.parametrize("kind", ["bond", "angle", "torsion"])
def test_bonded_hvp(ligand, kind):
    """Bonded terms with a lambda-dependent prefactor (mult 1, offset 0 on a
    third of the terms)."""
    rng = np.random.default_rng(15)
    s = ligand
    idxs, params = {"bond": (s.bond_idxs, s.bond_params), "angle": (s.angle_idxs, s.angle_params),
                    "torsion": (s.torsion_idxs, s.torsion_params)}[kind]
    n = len(idxs)
    mult = np.zeros(n, np.int32)
    offset = np.ones(n, np.int32)
    mult[::3] = 1
    offset[::3] = 0
    cls = {"bond": potentials.HarmonicBond, "angle": potentials.HarmonicAngle,
           "torsion": potentials.PeriodicTorsion}[kind]
    pot = cls(idxs, mult, offset).unbound_impl(F32)
    vx = rng.normal(size=(s.N, 3))
    vp = rng.normal(size=params.shape) * 0.1
    out = pot.execute_hvp(s.x, params, s.box, 0.3, vx, vp, 0.6)
    fn = obonded.bonded_energy_fn(kind, idxs, mult, offset)
    ref = onb.hvp(fn, s.x, params, 0.3, vx, vp, 0.6)
    hx, hp, hl = out
    ex = onb.force_rel_err(hx, ref[0]).max()
    ep = np.abs(hp - ref[1]).max() / max(np.abs(ref[1]).max(), 1.0)
    print(f"[{kind}] hx {ex:.3e} hp {ep:.3e} hl {hl:.9g} ref {ref[2]:.9g}", flush=True)
    assert ex <= TOL and ep <= TOL
    assert abs(hl - ref[2]) <= TOL * max(abs(ref[2]), 1.0)


def test_summed_hvp_equals_sum_of_parts(ligand):
    """SummedPotential splits the parameter tangent and sums the coordinate
    and lambda tangents: the sum of its parts."""
    rng = np.random.default_rng(16)
    s = ligand
    bps = s.bound_potentials()
    sp = potentials.SummedPotential(bps, [bp.params for bp in bps])
    impl = sp.unbound_impl(F32)
    flat = np.concatenate([np.asarray(bp.params, np.float64).reshape(-1) for bp in bps])
    vx = rng.normal(size=(s.N, 3))
    vps = [rng.normal(size=np.asarray(bp.params).shape) * 0.01 for bp in bps]
    for v, bp in zip(vps, bps):
        if v.ndim == 2 and v.shape[1] == 3 and v.shape[0] == s.N:
            v[np.asarray(bp.params)[:, 2] == 0.0, 2] = 0.0
    vflat = np.concatenate([v.reshape(-1) for v in vps])
    hx, hp, hl = impl.execute_hvp(s.x, flat, s.box, 0.4, vx, vflat, 0.3)
    sx, sl, sps = np.zeros_like(hx), 0.0, []
    for bp, v in zip(bps, vps):
        px, pp, pl = bp.unbound_impl(F32).execute_hvp(s.x, bp.params, s.box, 0.4, vx, v, 0.3)
        sx += px
        sl += pl
        sps.append(np.asarray(pp).reshape(-1))
    # each call scales its tangent by its own power of two (its largest entry
    # differs), so the sums agree to the fixed-point resolution of each call
    res = len(bps) * 2.0 ** -36 * max(np.abs(vx).max(), np.abs(vflat).max(), 0.3) * 2
    np.testing.assert_allclose(hx, sx, rtol=1e-14, atol=res)
    np.testing.assert_allclose(hp.reshape(-1), np.concatenate(sps), rtol=1e-14, atol=4 * res)
    assert abs(hl - sl) <= 1e-14 * abs(sl) + res


def test_hvp_bitwise_run_to_run_and_symmetric(ligand):
    rng = np.random.default_rng(17)
    s = ligand
    pot = s.nonbonded().unbound_impl(F32)
    u = (*_tangents(s, rng), 0.4)
    v = (*_tangents(s, rng), -0.9)
    a = pot.execute_hvp(s.x, s.nb_params, s.box, 0.5, *u)
    b = pot.execute_hvp(s.x, s.nb_params, s.box, 0.5, *u)
    for p, q in zip(a, b):
        np.testing.assert_array_equal(p, q)
    c = pot.execute_hvp(s.x, s.nb_params, s.box, 0.5, *v)
    vhu = sum(float(np.sum(np.asarray(p) * np.asarray(q))) for p, q in zip(v, a))
    uhv = sum(float(np.sum(np.asarray(p) * np.asarray(q))) for p, q in zip(u, c))
    assert abs(vhu - uhv) <= 1e-9 * max(abs(vhu), 1.0)


def test_hvp_of_zero_tangent_is_zero(ligand):
    s = ligand
    pot = s.nonbonded().unbound_impl(F32)
    hx, hp, hl = pot.execute_hvp(s.x, s.nb_params, s.box, 0.5)
    assert not np.any(hx) and not np.any(hp) and hl == 0.0


@pytest.fixture(scope="module")
def mutated(gpu):
    s = ts.mutate_ligand(ts.ligand_in_water(side=3.0, n_waters=800, n_heavy=12, n_ligand=22, seed=3))
    r = lambda p: p.astype(np.float32).astype(np.float64)
    return dataclasses.replace(s, nb_params=r(s.nb_params), nb_params_final=r(s.nb_params_final))


@pytest.fixture(scope="module")
def mutated_soft(mutated):
    """The parts alone (all pairs, or exclusions) need du/dp itself in range
    (the chain rule through p(lambda) uses it): at bonded ligand distances
    du/dsigma reaches ~9e7, beyond the 2^63 / 2^37 ~ 6.7e7 range of that
    column, where the full potential's two parts cancel. Ligand epsilons
    (both ends) x 0.1 keep it at ~1e6."""
    lig = np.asarray(mutated.ligand_idxs)
    p0, p1 = mutated.nb_params.copy(), mutated.nb_params_final.copy()
    for p in (p0, p1):
        p[lig, 2] = (p[lig, 2] * 0.1).astype(np.float32)
    return dataclasses.replace(mutated, nb_params=p0, nb_params_final=p1)


def _interpolated_f32(energy_fn):
    """onb.interpolated with p(lambda) valued at its FP32 rounding, the value
    the kernels evaluate; the derivatives are the exact interpolation's (the
    rounding enters as a constant offset)."""

    def fn(xt, pt, lt):
        n = pt.shape[0] // 2
        p = (1 - lt) * pt[:n] + lt * pt[n:]
        return energy_fn(xt, p + (p.detach().float().double() - p.detach()), lt)

    return fn


def _interp_tangents(s, rng, ligand_scale=1.0):
    """Tangents on (coords, initial params, final params); epsilon tangents
    stay 0 where epsilon is 0 at both ends (LJ off throughout)."""
    vx, vp0 = _tangents(s, rng, ligand_scale)
    _, vp1 = _tangents(s, rng, ligand_scale)
    off = (s.nb_params[:, 2] == 0.0) & (s.nb_params_final[:, 2] == 0.0)
    vp0[off, 2] = 0.0
    vp1[off, 2] = 0.0
    return vx, np.concatenate([vp0, vp1])


@pytest.mark.parametrize("lam", [0.35, 0.8])
def test_interpolated_all_pairs_hvp(mutated_soft, lam):
    """NonbondedAllPairs over [2N, 3] parameters p(lambda) = (1 - lambda) p0 +
    lambda p1 (the lambda tangent moves the ligand parameters too, so it is
    kept small like the ligand's: see _tangents)."""
    rng = np.random.default_rng(21)
    s = mutated_soft
    pot = potentials.NonbondedAllPairsInterpolated(s.lambda_plane_idxs, s.lambda_offset_idxs, s.beta,
                                                   s.cutoff).unbound_impl(F32)
    vx, vp = _interp_tangents(s, rng, ligand_scale=1e-4)
    out = pot.execute_hvp(s.x, s.nb_params_all, s.box, lam, vx, vp, 1e-4)
    fn = onb.pair_energy_fn(s.x, s.box, np.zeros((0, 2), int), np.zeros((0, 2)), s.beta, s.cutoff,
                            s.lambda_plane_idxs, s.lambda_offset_idxs)
    ref = onb.hvp(_interpolated_f32(fn), s.x, s.nb_params_all, lam, vx, vp, 1e-4)
    _check(out, ref, f"interpolated all pairs lam={lam}")


def test_interpolated_exclusions_hvp(mutated_soft):
    rng = np.random.default_rng(22)
    s = mutated_soft
    pot = potentials.NonbondedPairListNegatedInterpolated(s.exclusion_idxs, s.scales, s.lambda_plane_idxs,
                                                          s.lambda_offset_idxs, s.beta, s.cutoff).unbound_impl(F32)
    vx, vp = _interp_tangents(s, rng, ligand_scale=1e-4)
    out = pot.execute_hvp(s.x, s.nb_params_all, s.box, 0.6, vx, vp, -1e-4)
    fn = onb.pair_list_energy_fn(s.exclusion_idxs, s.scales, s.box, s.beta, s.cutoff, s.lambda_plane_idxs,
                                 s.lambda_offset_idxs, negated=True)
    ref = onb.hvp(_interpolated_f32(fn), s.x, s.nb_params_all, 0.6, vx, vp, -1e-4)
    _check(out, ref, "interpolated exclusions")


def test_interpolated_nonbonded_hvp(mutated):
    """The full interpolated Nonbonded potential with unscaled tangents: the
    all-pairs and exclusion terms of bonded ligand pairs leave the fixed-point
    range alone and cancel in the shared sums before the one chain rule
    through p(lambda)."""
    rng = np.random.default_rng(23)
    s = mutated
    pot = s.nonbonded().unbound_impl(F32)
    vx, vp = _interp_tangents(s, rng)
    out = pot.execute_hvp(s.x, s.nb_params_all, s.box, 0.45, vx, vp, 0.7)
    ref = onb.hvp(_interpolated_f32(_net_nonbonded_fn(s)), s.x, s.nb_params_all, 0.45, vx, vp, 0.7)
    _check(out, ref, "interpolated nonbonded")


def _net_nonbonded_fn(s):
    """Oracle Nonbonded energy without cancellation (see test_nonbonded_hvp)."""
    pairs = onb.candidate_pairs(s.x, s.box, s.cutoff)
    excluded = {tuple(p) for p in np.sort(np.asarray(s.exclusion_idxs), axis=1).tolist()}
    keep = np.array([tuple(p) not in excluded for p in pairs.tolist()], dtype=bool)
    fn_pairs = onb.pair_energy_fn(s.x, s.box, np.zeros((0, 2), int), np.zeros((0, 2)), s.beta, s.cutoff,
                                  s.lambda_plane_idxs, s.lambda_offset_idxs, pairs=pairs[keep])
    fn_rest = onb.pair_list_energy_fn(s.exclusion_idxs, 1.0 - np.asarray(s.scales), s.box, s.beta, s.cutoff,
                                      s.lambda_plane_idxs, s.lambda_offset_idxs)
    return lambda xt, pt, lt: fn_pairs(xt, pt, lt) + fn_rest(xt, pt, lt)


def test_force_matching_loss_backpropagates_to_parameters(ligand):
    """A loss on forces and du/dlambda, loss = |du/dx - F|^2 + (du/dl)^2, through
    the Torch interface (one SummedPotential custom op): its gradient in
    coordinates, every parameter array and lambda takes the double backward
    (Hessian-vector products). Checked against the same loss on the FP64
    oracle energy differentiated twice by Torch; relative 1e-5 of each
    gradient's largest entry (the forward forces carry the force kernels'
    ~1e-7 relative error)."""
    import torch

    from timemachine_amd.fe import functional

    s = ligand
    rng = np.random.default_rng(21)
    bps = s.bound_potentials()
    params = [np.asarray(bp.params, np.float64) for bp in bps]
    U = functional.construct_differentiable_interface_fast(bps, params)
    target = torch.tensor(rng.normal(size=(s.N, 3)) * 300.0, dtype=torch.float64)

    def grads(energy):
        x = torch.tensor(s.x, dtype=torch.float64, requires_grad=True)
        ps = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in params]
        lam = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
        gx, gl = torch.autograd.grad(energy(x, ps, lam), (x, lam), create_graph=True)
        loss = ((gx - target) ** 2).sum() + gl ** 2
        return torch.autograd.grad(loss, (x, *ps, lam))

    got = grads(lambda x, ps, lam: U(x, ps, s.box, lam))
    kinds = ["bond", "angle", "torsion"]
    idxs = [s.bond_idxs, s.angle_idxs, s.torsion_idxs]
    fns = [obonded.bonded_energy_fn(k, i) for k, i in zip(kinds, idxs)] + [_net_nonbonded_fn(s)]
    assert len(fns) == len(bps)
    want = grads(lambda x, ps, lam: sum(f(x, p, lam) for f, p in zip(fns, ps)))
    names = ["x", *kinds, "nonbonded", "lambda"]
    for name, a, b in zip(names, got, want):
        a = a.numpy()
        b = b.numpy()
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-12)
        print(f"[force matching] d loss / d {name}: max err / max|ref| = {err:.3e}", flush=True)
        assert err <= 1e-5
