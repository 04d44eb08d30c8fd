"""Parameter-interpolated nonbonded on the GPU (HIP path through the C ABI).

Parameters are [2N, 3] = (initial, final), used at lambda as
(1 - lambda) initial + lambda final (lib/potentials.py:347-368,
tests/test_parameter_interpolation.py:17-68 in the reference).

* Parity against the FP64 oracle on a ligand in water whose ligand changes
  charges, sigmas and epsilons (some ligand atoms lose their LJ term, others
  gain one) at lambda 0, 0.2, 0.6 and 1, with and without 4D offsets: forces
  |dF| / max(|F|, 1) <= 1e-5 per atom, |dU| <= 1e-6 |U|, |d du_dl| <=
  1e-5 max(|du_dl|, 1), du_dp [2N, 3] per column <= 1e-5 max|column|.
* Equal initial and final parameters give the plain potential's bits, and
  without offsets du_dl is exactly 0 (the known answer of
  tests/test_nonbonded.py:303-320).
* Tile kernel == pair-list kernel over the same pairs, bit for bit; Hilbert
  sort and padding do not change bits (tests/test_nonbonded.py:322-342).
* An excluded pair cancels exactly; MD with the interpolated potential is
  bitwise the same alone and inside a batch.
"""

import numpy as np
import pytest

from oracle import nonbonded as onb
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import LangevinIntegrator, custom_ops, potentials

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.fixture(scope="module")
def rbfe(gpu):
    return ts.mutate_ligand(ts.ligand_in_water(side=3.2, n_waters=1050, n_heavy=24, n_ligand=40, seed=5), seed=3)


@pytest.fixture(scope="module")
def rbfe_no_offsets(rbfe):
    return ts.mutate_ligand(ts.ligand_in_water(side=3.2, n_waters=1050, n_heavy=24, n_ligand=40, seed=5), seed=3,
                            offsets=False)


def _oracle(s, lam):
    return onb.nonbonded_pairs(s.x, s.nb_params_all, s.box, lam, s.exclusion_idxs, s.scales, s.beta, s.cutoff,
                               s.lambda_plane_idxs, s.lambda_offset_idxs, interpolated=True)


def _check(out, ref, label):
    du_dx, du_dp, du_dl, u = out
    rdx, rdp, rdl, ru = ref
    assert du_dp.shape == rdp.shape
    ferr = onb.force_rel_err(du_dx, rdx)
    print(f"[{label}] force rel err {ferr.max():.3e}  u {u:.6f} ref {ru:.6f}  du_dl {du_dl:.6f} ref {rdl:.6f}",
          flush=True)
    assert ferr.max() <= 1e-5
    assert abs(u - ru) <= 1e-6 * abs(ru)
    assert abs(du_dl - rdl) <= 1e-5 * max(abs(rdl), 1.0)
    for c in range(3):
        scale = max(np.abs(rdp[:, c]).max(), 1e-12)
        err = np.abs(du_dp[:, c] - rdp[:, c]).max() / scale
        print(f"[{label}] du_dp column {c} max err / max|ref| = {err:.3e}", flush=True)
        assert err <= 1e-5


@pytest.mark.parametrize("lam", [0.0, 0.2, 0.6, 1.0])
def test_parity_with_offsets(rbfe, lam):
    pot = rbfe.nonbonded().unbound_impl(F32)
    out = pot.execute(rbfe.x, rbfe.nb_params_all, rbfe.box, lam)
    _check(out, _oracle(rbfe, lam), f"offsets lam={lam}")
    n = rbfe.N
    if lam == 0.0:
        assert np.all(out[1][n:] == 0.0)  # lambda * g
    if lam == 1.0:
        assert np.all(out[1][:n] == 0.0)  # (1 - lambda) * g


@pytest.mark.parametrize("lam", [0.0, 0.6])
def test_parity_parameters_only(rbfe_no_offsets, lam):
    s = rbfe_no_offsets
    out = s.nonbonded().unbound_impl(F32).execute(s.x, s.nb_params_all, s.box, lam)
    _check(out, _oracle(s, lam), f"no offsets lam={lam}")
    assert out[2] != 0.0


def test_equal_endpoints_reduce_to_plain(rbfe, rbfe_no_offsets):
    for s in (rbfe, rbfe_no_offsets):
        plain = potentials.Nonbonded(s.exclusion_idxs, s.scales, s.lambda_plane_idxs, s.lambda_offset_idxs, s.beta,
                                     s.cutoff).unbound_impl(F32)
        interp = potentials.NonbondedInterpolated(s.exclusion_idxs, s.scales, s.lambda_plane_idxs,
                                                  s.lambda_offset_idxs, s.beta, s.cutoff).unbound_impl(F32)
        P = np.concatenate([s.nb_params, s.nb_params])
        for lam in (0.0, 0.41):
            a = plain.execute_fixed_raw(s.x, s.nb_params, s.box, lam)
            b = interp.execute_fixed_raw(s.x, P, s.box, lam)
            assert np.array_equal(a[0], b[0])  # du_dx
            assert np.sum(a[2], dtype=np.uint64) == np.sum(b[2], dtype=np.uint64)  # du_dl
            assert np.sum(a[3], dtype=np.uint64) == np.sum(b[3], dtype=np.uint64)  # u
    s = rbfe_no_offsets
    P = np.concatenate([s.nb_params, s.nb_params])
    pot = potentials.NonbondedInterpolated(s.exclusion_idxs, s.scales, s.lambda_plane_idxs, s.lambda_offset_idxs,
                                           s.beta, s.cutoff).unbound_impl(F32)
    for lam in (0.0, 0.37, 1.0):
        assert pot.execute(s.x, P, s.box, lam)[2] == 0.0


@pytest.mark.parametrize("lam", [0.0, 0.45])
def test_tile_kernel_equals_pair_list_bitwise(rbfe, lam):
    s = rbfe
    pairs = onb.candidate_pairs(s.x, s.box, s.cutoff).astype(np.int32)
    ones = np.ones((len(pairs), 2))
    allp = custom_ops.NonbondedAllPairs_f32_interpolated(None, s.lambda_plane_idxs, s.lambda_offset_idxs, s.beta,
                                                         s.cutoff)
    plist = custom_ops.NonbondedPairList_f32_interpolated(None, pairs, ones, s.lambda_plane_idxs,
                                                          s.lambda_offset_idxs, s.beta, s.cutoff)
    a = allp.execute_fixed_raw(s.x, s.nb_params_all, s.box, lam)
    b = plist.execute_fixed_raw(s.x, s.nb_params_all, s.box, lam)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.sum(a[2], dtype=np.uint64) == np.sum(b[2], dtype=np.uint64)
    assert np.sum(a[3], dtype=np.uint64) == np.sum(b[3], dtype=np.uint64)


def test_sort_and_padding_do_not_change_bits(rbfe):
    base = None
    for sort in (True, False):
        for pad in (0.0, 0.2):
            pot = rbfe.nonbonded().unbound_impl(F32)
            if not sort:
                pot.disable_hilbert_sort()
            pot.set_nblist_padding(pad)
            res = pot.execute_fixed_raw(rbfe.x, rbfe.nb_params_all, rbfe.box, 0.3)
            cur = [res[0], res[1], np.sum(res[2], dtype=np.uint64), np.sum(res[3], dtype=np.uint64)]
            if base is None:
                base = cur
            else:
                for p, q in zip(base, cur):
                    np.testing.assert_array_equal(p, q)


def test_exclusion_cancels_exactly(gpu):
    x = np.array([[0.1, 0.2, 0.3], [0.35, 0.3, 0.25], [1.6, 1.6, 1.6]])
    p0 = np.array([[3.0, 0.15, 0.8], [-2.0, 0.12, 0.0], [0.0, 0.0, 0.0]])
    p1 = np.array([[1.0, 0.11, 0.3], [-1.0, 0.14, 0.9], [0.0, 0.0, 0.0]])
    box = np.eye(3) * 3.0
    nb = potentials.NonbondedInterpolated(np.array([[0, 1]], np.int32), np.array([[1.0, 1.0]]),
                                          np.zeros(3, np.int32), np.array([0, 1, 0], np.int32), 2.0,
                                          1.2).unbound_impl(F32)
    for lam in (0.0, 0.3, 1.0):
        du_dx, du_dp, du_dl, u = nb.execute_fixed_raw(x, np.concatenate([p0, p1]), box, lam)
        assert np.all(du_dx == 0) and np.all(du_dp == 0)
        assert np.sum(u, dtype=np.uint64) == 0 and np.sum(du_dl, dtype=np.uint64) == 0


def test_sizes_and_unsupported_modes(rbfe):
    pot = rbfe.nonbonded().unbound_impl(F32)
    with pytest.raises(RuntimeError, match="expected P"):
        pot.execute(rbfe.x, rbfe.nb_params, rbfe.box, 0.5)  # [N, 3] where [2N, 3] is needed
    with pytest.raises(RuntimeError):
        pot.execute_dl_dp(rbfe.x, rbfe.nb_params_all, rbfe.box, 0.5)
    with pytest.raises(RuntimeError, match="transform"):
        custom_ops.NonbondedAllPairs_f32_interpolated(None, rbfe.lambda_plane_idxs, rbfe.lambda_offset_idxs, 2.0,
                                                      1.2, None, "lambda*lambda")


def _ctx(s, lam_seed):
    bps = [p.bound_impl(F32) for p in s.bound_potentials()]
    intg = LangevinIntegrator(300.0, 0.0025, 1.0, s.masses, lam_seed).impl()
    v0 = ts.maxwell_boltzmann(s.masses, 300.0, lam_seed)
    return custom_ops.Context(s.x, v0, s.box, intg, bps), bps


def test_md_alone_equals_batched(rbfe):
    """Three windows of the interpolated system: each run alone and all three in
    one batch give the same coordinates and du_dl series, bit for bit."""
    lams = [0.1, 0.5, 0.9]
    alone = []
    for w, lam in enumerate(lams):
        ctx, keep = _ctx(rbfe, 100 + w)
        du_dl, xs, _ = ctx.multiple_steps(np.full(40, lam), 10, 40)
        alone.append((du_dl, ctx.get_x_t()))
        del ctx, keep
    ctxs, keeps = zip(*[_ctx(rbfe, 100 + w) for w in range(3)])
    batch = custom_ops.BatchedContexts(list(ctxs), lams)
    du_dl, _, _ = batch.multiple_steps(40, store_du_dl_interval=10)
    for w in range(3):
        np.testing.assert_array_equal(du_dl[w], alone[w][0])
        np.testing.assert_array_equal(batch.get_x_t(w), alone[w][1])
    assert np.all(np.isfinite(du_dl)) and np.any(du_dl != 0.0)


def test_batch_rejects_mixed_interpolation(rbfe):
    plain = ts.ligand_in_water(side=3.2, n_waters=1050, n_heavy=24, n_ligand=40, seed=5)
    (a, ka), (b, kb) = _ctx(rbfe, 1), _ctx(plain, 2)
    with pytest.raises(RuntimeError, match="interpolated"):
        custom_ops.BatchedContexts([a, b], [0.5, 0.5])
    batch = custom_ops.BatchedContexts([a], [0.5])
    with pytest.raises(RuntimeError, match="interpolated"):
        batch.set_grad_observer(5)
