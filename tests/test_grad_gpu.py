"""d(du/dlambda)/dp on the GPU (HIP path through the C ABI) against the FP64
oracle's double backward.

Tolerance: per parameter column, max |d| <= 1e-5 * max |column of the oracle|
(the du_dp tolerance; 1e-9 plus one fixed-point quantum 2^-36 for the FP64
bonded terms, each parameter belonging to one term), and exact zeros where
the mixed derivative vanishes identically (no offsets, lambda = 0, mult = 0).
"""

import numpy as np
import pytest

from oracle import bonded as obonded
from oracle import nonbonded as onb
from timemachine_amd import testsystems as ts
from timemachine_amd.lib import potentials

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.fixture(scope="module")
def ligand(gpu):
    return ts.ligand_in_water()


def _alchemical_pairs(sys):
    """Candidate pairs with at least one atom carrying a lambda offset; every
    other pair has dw = 0 and no lambda dependence."""
    pairs = onb.candidate_pairs(sys.x, sys.box, sys.cutoff)
    alch = sys.lambda_offset_idxs != 0
    return pairs[alch[pairs[:, 0]] | alch[pairs[:, 1]]]


FIXED_QUANTUM = 2.0**-36  # resolution of the 64-bit fixed-point accumulators


def _check_columns(got, ref, label, tol=1e-5, atol=0.0):
    """max |got - ref| <= tol * max|ref| + atol per column; atol covers the
    fixed-point rounding (one quantum per accumulated term)."""
    for c in range(ref.shape[1]):
        scale = np.abs(ref[:, c]).max()
        if scale == 0.0:
            assert np.all(got[:, c] == 0.0), label
            continue
        err = np.abs(got[:, c] - ref[:, c]).max()
        print(f"[{label}] column {c}: max err {err:.3e}, max|ref| {scale:.3e}, rel {err / scale:.3e}")
        assert err <= tol * scale + atol


@pytest.mark.parametrize("lam", [0.0, 0.3, 0.5, 0.9])
def test_nonbonded_dl_dp_parity(ligand, lam):
    pot = ligand.nonbonded().unbound_impl(F32)
    got = pot.execute_dl_dp(ligand.x, ligand.nb_params, ligand.box, lam)
    ref = onb.nonbonded_pairs_dl_dp(
        ligand.x, ligand.nb_params, ligand.box, lam, ligand.exclusion_idxs, ligand.scales, ligand.beta,
        ligand.cutoff, ligand.lambda_plane_idxs, ligand.lambda_offset_idxs, pairs=_alchemical_pairs(ligand),
    )
    if lam == 0.0:
        # w = 0 for every atom: du/dlambda and its parameter derivatives vanish
        assert np.all(ref == 0.0)
    _check_columns(got, ref, f"nonbonded dl_dp lam={lam}")
    water = ligand.lambda_offset_idxs == 0
    # water-water pairs do not depend on lambda; only rows near the ligand can be nonzero
    assert np.count_nonzero(np.abs(got[water]).sum(1)) < water.sum()


def test_nonbonded_dl_dp_bitwise_repeatable(ligand):
    pot = ligand.nonbonded().unbound_impl(F32)
    a = pot.execute_dl_dp(ligand.x, ligand.nb_params, ligand.box, 0.5)
    b = pot.execute_dl_dp(ligand.x, ligand.nb_params, ligand.box, 0.5)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["bond", "angle", "torsion"])
def test_bonded_dl_dp_parity(ligand, kind):
    """Ligand terms bound with lambda multipliers (mult +1 / -1 / 2 alternating)."""
    idxs = {"bond": ligand.bond_idxs, "angle": ligand.angle_idxs, "torsion": ligand.torsion_idxs}[kind]
    params = {"bond": ligand.bond_params, "angle": ligand.angle_params, "torsion": ligand.torsion_params}[kind]
    lig = set(np.asarray(ligand.ligand_idxs).tolist())
    keep = np.array([all(int(a) in lig for a in row) for row in idxs])
    idxs, params = np.ascontiguousarray(idxs[keep]), np.ascontiguousarray(params[keep])
    mult = np.array([1, -1, 2], np.int32)[np.arange(len(idxs)) % 3]
    off = np.ones(len(idxs), np.int32)
    cls = {"bond": potentials.HarmonicBond, "angle": potentials.HarmonicAngle, "torsion": potentials.PeriodicTorsion}[
        kind
    ]
    pot = cls(idxs, mult, off).unbound_impl(F32)
    got = pot.execute_dl_dp(ligand.x, params, ligand.box, 0.4)
    ref = obonded.bonded_dl_dp(kind, ligand.x, params, 0.4, idxs, mult, off)
    _check_columns(got, ref, f"{kind} dl_dp", tol=1e-9, atol=FIXED_QUANTUM)
    # mult = 0 terms have no lambda dependence
    pot0 = cls(idxs).unbound_impl(F32)
    assert np.all(pot0.execute_dl_dp(ligand.x, params, ligand.box, 0.4) == 0.0)
