"""Nonbonded parity on the GPU (HIP path through the C ABI) against the FP64 oracle.

Tolerances (stated targets): forces |dF| / max(|F|, 1) <= 1e-5 per atom,
energy |dU| / |U| <= 1e-6, du_dl |d| <= 1e-5 * max(|du_dl|, 1), du_dp per
column |d| <= 1e-5 * max|column|. Fixed-point outputs must be bitwise
identical run to run, across sort orders, paddings, output selections and
batch composition.
"""

import numpy as np
import pytest

from oracle import nblist as onblist
from oracle import nonbonded as onb
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import custom_ops, potentials

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.fixture(scope="module")
def water(gpu):
    return ts.water_box()


@pytest.fixture(scope="module")
def ligand(gpu):
    return ts.ligand_in_water()


def _oracle(sys, lam):
    return onb.nonbonded_pairs(
        sys.x, sys.nb_params, sys.box, lam, sys.exclusion_idxs, sys.scales, sys.beta, sys.cutoff,
        sys.lambda_plane_idxs, sys.lambda_offset_idxs,
    )


def _check_parity(out, ref, label):
    du_dx, du_dp, du_dl, u = out
    rdx, rdp, rdl, ru = ref
    ferr = onb.force_rel_err(du_dx, rdx)
    print(f"[{label}] max force rel err {ferr.max():.3e}  u {u:.6f} ref {ru:.6f} rel {abs(u - ru) / abs(ru):.3e}  "
          f"du_dl {du_dl:.6f} ref {rdl:.6f}")
    assert ferr.max() <= 1e-5
    assert abs(u - ru) <= 1e-6 * abs(ru)
    assert abs(du_dl - rdl) <= 1e-5 * max(abs(rdl), 1.0)
    for c in range(3):
        scale = max(np.abs(rdp[:, c]).max(), 1e-12)
        err = np.abs(du_dp[:, c] - rdp[:, c]).max() / scale
        print(f"[{label}] du_dp column {c} max err / max|ref| = {err:.3e}")
        assert err <= 1e-5


def test_water_box_parity(water):
    """Water box, 887 waters, 3.0 nm, rc 1.2, beta 2, lambda 0: all four outputs vs FP64."""
    pot = water.nonbonded().unbound_impl(F32)
    out = pot.execute(water.x, water.nb_params, water.box, 0.0)
    _check_parity(out, _oracle(water, 0.0), "water")
    assert out[2] == 0.0  # every tile is vanilla


@pytest.mark.parametrize("lam", [0.0, 0.35, 0.5, 1.0])
def test_ligand_parity(ligand, lam):
    """Ligand in water with 4D offsets on the ligand: decoupling tiles and du_dl."""
    pot = ligand.nonbonded().unbound_impl(F32)
    out = pot.execute(ligand.x, ligand.nb_params, ligand.box, lam)
    _check_parity(out, _oracle(ligand, lam), f"ligand lam={lam}")


def test_run_to_run_bitwise(water):
    pot = water.nonbonded().unbound_impl(F32)
    a = pot.execute_fixed_raw(water.x, water.nb_params, water.box, 0.0)
    b = pot.execute_fixed_raw(water.x, water.nb_params, water.box, 0.0)
    for p, q in zip(a, b):
        np.testing.assert_array_equal(p, q)


def test_sort_and_padding_do_not_change_bits(ligand):
    """Hilbert sort on/off and any neighborlist padding give identical fixed-point
    results (tests/test_nonbonded.py:48-158 in the reference)."""
    base = None
    for sort in (True, False):
        for pad in (0.0, 0.1, 0.25):
            pot = ligand.nonbonded().unbound_impl(F32)
            if not sort:
                pot.disable_hilbert_sort()
            pot.set_nblist_padding(pad)
            res = pot.execute_fixed_raw(ligand.x, ligand.nb_params, ligand.box, 0.4)
            sums = [np.sum(r, dtype=np.uint64) for r in res[2:]]
            cur = [res[0], res[1], *sums]
            if base is None:
                base = cur
            else:
                for p, q in zip(base, cur):
                    np.testing.assert_array_equal(p, q)


def test_newton_third_law_exact(ligand):
    """Every pair adds +g to one atom and -g to the other: the fixed-point sum is 0."""
    pot = ligand.nonbonded().unbound_impl(F32)
    du_dx = pot.execute_fixed_raw(ligand.x, ligand.nb_params, ligand.box, 0.3)[0].reshape(-1, 3)
    assert np.all(np.sum(du_dx, axis=0, dtype=np.uint64) == 0)


def test_exclusion_cancels_exactly(gpu):
    """AllPairs + PairListNegated(scale 1) on an excluded pair is exactly zero."""
    x = np.array([[0.1, 0.2, 0.3], [0.35, 0.3, 0.25], [1.6, 1.6, 1.6]])
    p = np.array([[3.0, 0.15, 0.8], [-2.0, 0.12, 0.9], [0.0, 0.0, 0.0]])
    box = np.eye(3) * 3.0
    nb = potentials.Nonbonded(np.array([[0, 1]], np.int32), np.array([[1.0, 1.0]]), np.zeros(3, np.int32),
                              np.array([0, 1, 0], np.int32), 2.0, 1.2).unbound_impl(F32)
    for lam in (0.0, 0.3):
        du_dx, du_dp, du_dl, u = nb.execute_fixed_raw(x, p, box, lam)
        assert np.all(du_dx == 0) and np.all(du_dp == 0)
        assert np.sum(u, dtype=np.uint64) == 0 and np.sum(du_dl, dtype=np.uint64) == 0


@pytest.mark.parametrize("lam", [0.0, 0.5])
def test_tile_kernel_equals_pair_list_bitwise(ligand, lam):
    """The all-pairs tile kernel and the one-thread-per-pair list kernel use the
    same per-pair arithmetic, and fixed-point sums are exact: over the same set
    of pairs (every pair within the 3D cutoff, a superset of the 4D set) all
    four outputs must agree bit for bit."""
    pairs = onb.candidate_pairs(ligand.x, ligand.box, ligand.cutoff).astype(np.int32)
    ones = np.ones((len(pairs), 2))
    allp = custom_ops.NonbondedAllPairs_f32(None, ligand.lambda_plane_idxs, ligand.lambda_offset_idxs, ligand.beta,
                                            ligand.cutoff)
    plist = custom_ops.NonbondedPairList_f32(None, pairs, ones, ligand.lambda_plane_idxs, ligand.lambda_offset_idxs,
                                             ligand.beta, ligand.cutoff)
    a = allp.execute_fixed_raw(ligand.x, ligand.nb_params, ligand.box, lam)
    b = plist.execute_fixed_raw(ligand.x, ligand.nb_params, ligand.box, lam)
    assert np.array_equal(a[0], b[0])  # du_dx
    assert np.array_equal(a[1], b[1])  # du_dp
    assert np.sum(a[2], dtype=np.uint64) == np.sum(b[2], dtype=np.uint64)  # du_dl
    assert np.sum(a[3], dtype=np.uint64) == np.sum(b[3], dtype=np.uint64)  # u
    ga = allp.execute_dl_dp(ligand.x, ligand.nb_params, ligand.box, lam)
    gb = plist.execute_dl_dp(ligand.x, ligand.nb_params, ligand.box, lam)
    assert np.array_equal(ga, gb)


@pytest.mark.parametrize("n", [40, 100, 150, 250, 300, 701])
def test_tile_kernel_equals_pair_list_few_blocks(water, n):
    """1 to 11 row blocks, some atoms alchemical: every pair within the cutoff lands
    in exactly one tile (the diagonal blocks keep i < j only), so the tile kernel
    matches the pair list over all candidate pairs bit for bit."""
    x = water.x[:n]
    params = water.nb_params[:n]
    plane = np.zeros(n, np.int32)
    offset = (np.arange(n) % 7 == 0).astype(np.int32)
    pairs = onb.candidate_pairs(x, water.box, water.cutoff).astype(np.int32)
    ones = np.ones((len(pairs), 2))
    allp = custom_ops.NonbondedAllPairs_f32(None, plane, offset, water.beta, water.cutoff)
    plist = custom_ops.NonbondedPairList_f32(None, pairs, ones, plane, offset, water.beta, water.cutoff)
    for lam in (0.0, 0.3):
        a = allp.execute_fixed_raw(x, params, water.box, lam)
        b = plist.execute_fixed_raw(x, params, water.box, lam)
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1])
        assert np.sum(a[2], dtype=np.uint64) == np.sum(b[2], dtype=np.uint64)
        assert np.sum(a[3], dtype=np.uint64) == np.sum(b[3], dtype=np.uint64)


def test_selective_outputs_are_bitwise_consistent(ligand):
    """All 16 output selections give the same bits for what they compute."""
    pot = ligand.nonbonded().unbound_impl(F32)
    full = pot.execute_selective(ligand.x, ligand.nb_params, ligand.box, 0.25, True, True, True, True)
    for mask in range(16):
        flags = [bool(mask & (1 << k)) for k in range(4)]
        res = pot.execute_selective(ligand.x, ligand.nb_params, ligand.box, 0.25, *flags)
        for k in range(4):
            if flags[k]:
                np.testing.assert_array_equal(np.asarray(res[k]), np.asarray(full[k]))
            else:
                assert res[k] is None


def test_batch_matches_single(water):
    """execute_selective_batch over coords x params x lambdas == per-item execute, bitwise."""
    pot = water.nonbonded().unbound_impl(F32)
    rng = np.random.default_rng(0)
    xs = np.stack([water.x, (water.x + rng.normal(0, 0.002, water.x.shape)).astype(F32).astype(np.float64)])
    boxes = np.stack([water.box, water.box])
    ps = np.stack([water.nb_params, water.nb_params * np.array([1.0, 1.01, 0.99])])
    lams = np.array([0.0, 0.7])
    dx, dp, dl, u = pot.execute_selective_batch(xs, ps, boxes, lams, True, True, True, True)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                r = pot.execute(xs[i], ps[j], boxes[i], lams[k])
                np.testing.assert_array_equal(dx[i, j, k], r[0])
                np.testing.assert_array_equal(dp[i, j, k], r[1])
                assert dl[i, j, k] == r[2] and u[i, j, k] == r[3]


def test_box_validation_messages(water):
    pot = water.nonbonded().unbound_impl(F32)
    with pytest.raises(RuntimeError, match="Cutoff with padding is more than half of the box width"):
        pot.execute(water.x, water.nb_params, np.eye(3) * 2.0, 0.0)
    bad = water.box.copy()
    bad[0, 1] = 0.1
    with pytest.raises(RuntimeError, match="Provided non-ortholinear box"):
        pot.execute(water.x, water.nb_params, bad, 0.0)
    ap = custom_ops.NonbondedAllPairs_f32("", np.zeros(water.N, np.int32), np.zeros(water.N, np.int32), 2.0, 1.2)
    with pytest.raises(RuntimeError, match=r"expected N == N_, got N=3, N_=2661"):
        ap.execute(np.zeros((3, 3)), np.zeros((3, 3)), water.box, 0.0)
    with pytest.raises(RuntimeError, match=r"expected P == M\*N_\*3"):
        ap.execute(water.x, np.zeros((2, 3)), water.box, 0.0)


@pytest.mark.parametrize("size", [35, 64, 129, 1025, 1259, 2029])
def test_neighborlist_exact(gpu, size):
    """GPU tile neighborlist == the FP32 predicate, bit for bit (64-atom blocks), and
    == the f64 reference of tests/test_nblist.py:70-94 away from the cutoff boundary."""
    from pathlib import Path

    water = np.load(Path(__file__).parent / "golden" / "water.npy", allow_pickle=False)[:, :3]
    water = water.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(1234)
    x = water[rng.choice(np.arange(size), size, replace=False)]
    box = np.eye(3) * (x.max(0) - x.min(0) + 0.1)
    cutoff = 1.0
    nl = custom_ops.Neighborlist_f32(size)
    for _ in range(2):
        test = nl.get_nblist(x, box, cutoff, block_size=64)
        ref32 = onblist.fp32_ixn_list(x, box, cutoff, 64)
        assert [sorted(t) for t in test] == ref32
    ref64 = onblist.build_reference_ixn_list(x, box, cutoff, 64)
    nmis = sum(len(set(a) ^ set(b)) for a, b in zip(ref64, ref32))
    assert nmis <= 2
    # the reference's 32-atom row blocks (its default)
    test = nl.get_nblist(x, box, cutoff)
    fp32_blocks32 = onblist.fp32_ixn_list(x, box, cutoff, 32)
    assert test == fp32_blocks32
    ref_blocks32 = onblist.build_reference_ixn_list(x, box, cutoff, 32)
    assert sum(len(set(a) ^ set(b)) for a, b in zip(ref_blocks32, fp32_blocks32)) <= 2


def _morton_order(x, box, bits=5):
    """Atoms ordered along a Morton curve, so 64 consecutive atoms are compact."""
    L = np.diag(box)
    c = np.minimum((np.mod(x / L, 1.0) * (1 << bits)).astype(np.int64), (1 << bits) - 1)
    key = np.zeros(len(x), dtype=np.int64)
    for b in range(bits):
        for a in range(3):
            key |= ((c[:, a] >> b) & 1) << (3 * b + a)
    return np.argsort(key, kind="stable")


@pytest.mark.parametrize("cutoff", [1.0, 1.3, 1.9])
@pytest.mark.parametrize("order", ["morton", "given"])
def test_neighborlist_exact_full_box(ligand, cutoff, order):
    """6372 atoms in a 4 nm box, rows as compact blocks (Morton order) or as the
    system lists them (rows of the water lattice, wide blocks): the tile set is the
    FP32 predicate's, bit for bit, at the cutoff and well beyond it."""
    x = ligand.x[_morton_order(ligand.x, ligand.box)] if order == "morton" else ligand.x
    nl = custom_ops.Neighborlist_f32(ligand.N)
    test = nl.get_nblist(x, ligand.box, cutoff, block_size=64)
    ref32 = onblist.fp32_ixn_list(x, ligand.box, cutoff, 64)
    assert [sorted(t) for t in test] == ref32
    assert nl.get_nblist(x, ligand.box, cutoff) == onblist.fp32_ixn_list(x, ligand.box, cutoff, 32)


@pytest.mark.parametrize("pieces", [1, 2, 8])
def test_neighborlist_pieces_exact_full_box(ligand, pieces):
    """The build's waves per row block (1: large batches; 8: one system alone)
    do not change the tile set: the 6372-atom box at 1.3 nm equals the FP32
    predicate bit for bit with each split."""
    x = ligand.x[_morton_order(ligand.x, ligand.box)]
    nl = custom_ops.Neighborlist_f32(ligand.N)
    nl.set_pieces(pieces)
    test = nl.get_nblist(x, ligand.box, 1.3, block_size=64)
    ref32 = onblist.fp32_ixn_list(x, ligand.box, 1.3, 64)
    assert [sorted(t) for t in test] == ref32
    with pytest.raises(RuntimeError, match="pieces must be"):
        nl.set_pieces(3)


def test_neighborlist_resize_errors(gpu):
    nl = custom_ops.Neighborlist_f32(3)
    with pytest.raises(RuntimeError, match="size is must be at least 1"):
        nl.resize(0)
    with pytest.raises(RuntimeError, match="size is greater than max size"):
        nl.resize(4)


def test_neighborlist_covers_all_pairs_within_cutoff(ligand):
    """Every pair i < j within the cutoff has j in the list of i's row block."""
    nl = custom_ops.Neighborlist_f32(ligand.N)
    lists = nl.get_nblist(ligand.x, ligand.box, 1.3, block_size=64)
    M = np.zeros((len(lists), ligand.N), dtype=bool)
    for r, lst in enumerate(lists):
        M[r, lst] = True
    pairs = onb.candidate_pairs(ligand.x, ligand.box, 1.2)
    assert M[pairs[:, 0] // 64, pairs[:, 1]].all()
