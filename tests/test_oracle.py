"""Pins the oracle (CPU, no GPU): closed forms, brute force, finite differences,
invariants and the reference's own fixture data. A plausible mistake in the
oracle (dropped term, wrong sign, transposed index) fails one of these."""

import math
from pathlib import Path

import numpy as np
import pytest
import torch
from scipy.special import erfc

from oracle import barostat, bonded, integrator, nblist, nonbonded

GOLDEN = Path(__file__).resolve().parent / "golden"


def _tiny_system(rng, N=24, L=2.7, lam=0.3, with_offsets=True):
    x = rng.uniform(0, L, size=(N, 3))
    # keep atoms apart so energies are moderate
    params = np.stack([rng.normal(0, 3, N), rng.uniform(0.08, 0.16, N), rng.uniform(0.0, 1.0, N)], 1)
    params[::5, 2] = 0.0  # some eps = 0 atoms (LJ off)
    box = np.eye(3) * L
    excl = np.array([[0, 1], [2, 3], [4, 7], [10, 11]])
    scales = np.array([[1.0, 1.0], [0.5, 0.5], [1.0, 0.0], [0.3, 0.8]])
    plane = np.zeros(N, np.int32)
    offset = np.zeros(N, np.int32)
    if with_offsets:
        offset[:6] = 1
    return x, params, box, lam, excl, scales, plane, offset


def test_two_particle_closed_form():
    """u = q1 q2 erfc(beta r)/r + 4 eps ((s/r)^12 - (s/r)^6) with s = s1 + s2, eps = e1 e2."""
    r = 0.37
    x = np.array([[0.2, 0.3, 0.4], [0.2 + r, 0.3, 0.4]])
    p = np.array([[1.3, 0.12, 0.8], [-0.7, 0.14, 0.6]])
    box = np.eye(3) * 3.0
    beta, cutoff = 2.0, 1.2
    gx, gp, gl, u = nonbonded.nonbonded_v3(x, p, box, 0.0, np.zeros((0, 2), int), np.zeros((0, 2)), beta, cutoff, [0, 0], [0, 0])
    s, e = 0.26, 0.48
    u_ref = 1.3 * -0.7 * erfc(beta * r) / r + 4 * e * ((s / r) ** 12 - (s / r) ** 6)
    du_dr = 1.3 * -0.7 * (-2 * beta / math.sqrt(math.pi) * math.exp(-((beta * r) ** 2)) / r - erfc(beta * r) / r**2)
    du_dr += 4 * e * (-12 * s**12 / r**13 + 6 * s**6 / r**7)
    assert u == pytest.approx(u_ref, rel=1e-12)
    assert gx[1, 0] == pytest.approx(du_dr, rel=1e-10)
    assert gx[0, 0] == pytest.approx(-du_dr, rel=1e-10)
    # d/d q1 = q2 erfc/r ; d/d eps1 = e2 * 4((s/r)^12 - (s/r)^6) ; d/d sig1 = d/ds
    assert gp[0, 0] == pytest.approx(-0.7 * erfc(beta * r) / r, rel=1e-12)
    assert gp[0, 2] == pytest.approx(0.6 * 4 * ((s / r) ** 12 - (s / r) ** 6), rel=1e-10)
    assert gp[0, 1] == pytest.approx(4 * e * (12 * s**11 / r**12 - 6 * s**5 / r**6), rel=1e-10)
    assert gl == 0.0


def test_cutoff_and_4d_decoupling():
    """Beyond the cutoff nothing interacts; w = lambda * cutoff on one atom turns the pair off at lambda = 1."""
    x = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0]])
    p = np.array([[1.0, 0.1, 0.5], [1.0, 0.1, 0.5]])
    box = np.eye(3) * 3.0
    E = np.zeros((0, 2), int), np.zeros((0, 2))
    u_on = nonbonded.nonbonded_v3(x, p, box, 0.0, *E, 2.0, 1.2, [0, 0], [0, 1])[3]
    u_off = nonbonded.nonbonded_v3(x, p, box, 1.0, *E, 2.0, 1.2, [0, 0], [0, 1])[3]
    u_half = nonbonded.nonbonded_v3(x, p, box, 0.5, *E, 2.0, 1.2, [0, 0], [0, 1])[3]
    r = math.sqrt(0.25 + 0.6**2)
    assert u_off == 0.0
    assert u_on != 0.0
    s6 = (0.2 / r) ** 6
    assert u_half == pytest.approx(erfc(2 * r) / r + 4 * 0.25 * (s6 * s6 - s6), rel=1e-12)


def test_dense_equals_pair_evaluator(rng):
    args = _tiny_system(rng)
    a = nonbonded.nonbonded_v3(*args[:4], args[4], args[5], 2.0, 1.2, args[6], args[7])
    b = nonbonded.nonbonded_pairs(*args[:4], args[4], args[5], 2.0, 1.2, args[6], args[7])
    np.testing.assert_allclose(a[0], b[0], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-10, atol=1e-10)
    assert a[2] == pytest.approx(b[2], rel=1e-10, abs=1e-10)
    assert a[3] == pytest.approx(b[3], rel=1e-12)


def test_finite_differences(rng):
    x, p, box, lam, excl, sc, pl, of = _tiny_system(rng, N=12)
    gx, gp, gl, u = nonbonded.nonbonded_v3(x, p, box, lam, excl, sc, 2.0, 1.2, pl, of)
    qm, ljm = nonbonded.convert_exclusions_to_rescale_masks(excl, sc, len(x))

    def U(xx, pp, ll):
        T = lambda a: torch.tensor(a, dtype=torch.float64)  # noqa: E731
        return float(nonbonded.nonbonded_v3_energy(T(xx), T(pp), box, T(ll), qm, ljm, 2.0, 1.2, pl, of))

    h = 1e-6
    for i, k in [(0, 0), (3, 2), (7, 1), (11, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        assert gx[i, k] == pytest.approx((U(xp, p, lam) - U(xm, p, lam)) / (2 * h), rel=1e-5, abs=1e-5)
    for i, k in [(1, 0), (2, 1), (4, 2)]:
        pp, pm = p.copy(), p.copy()
        pp[i, k] += h
        pm[i, k] -= h
        assert gp[i, k] == pytest.approx((U(x, pp, lam) - U(x, pm, lam)) / (2 * h), rel=1e-5, abs=1e-5)
    assert gl == pytest.approx((U(x, p, lam + h) - U(x, p, lam - h)) / (2 * h), rel=1e-5, abs=1e-5)


def test_invariants(rng):
    x, p, box, lam, excl, sc, pl, of = _tiny_system(rng)
    gx, _, _, u = nonbonded.nonbonded_pairs(x, p, box, lam, excl, sc, 2.0, 1.2, pl, of)
    np.testing.assert_allclose(gx.sum(0), 0.0, atol=1e-9)  # Newton's third law
    u2 = nonbonded.nonbonded_pairs(x + np.array([0.31, -1.7, 2.9]), p, box, lam, excl, sc, 2.0, 1.2, pl, of)[3]
    assert u2 == pytest.approx(u, rel=1e-12)  # periodic translation invariance
    gl0 = nonbonded.nonbonded_pairs(x, p, box, lam, excl, sc, 2.0, 1.2, pl, np.zeros_like(of))[2]
    assert gl0 == 0.0  # no offsets -> no lambda dependence


def test_full_exclusion_removes_pair(rng):
    x, p, box, lam, _, _, pl, of = _tiny_system(rng, with_offsets=False)
    full = nonbonded.nonbonded_pairs(x, p, box, 0.0, np.array([[2, 9]]), np.array([[1.0, 1.0]]), 2.0, 1.2, pl, of)[3]
    none = nonbonded.nonbonded_pairs(x, p, box, 0.0, np.zeros((0, 2), int), np.zeros((0, 2)), 2.0, 1.2, pl, of)[3]
    d = x[2] - x[9]
    d -= np.diag(box) * np.floor(d / np.diag(box) + 0.5)
    r = np.linalg.norm(d)
    u_pair = 0.0
    if r < 1.2:
        u_pair = p[2, 0] * p[9, 0] * erfc(2 * r) / r
        if p[2, 2] != 0 and p[9, 2] != 0:
            s = p[2, 1] + p[9, 1]
            u_pair += 4 * p[2, 2] * p[9, 2] * ((s / r) ** 12 - (s / r) ** 6)
    assert full == pytest.approx(none - u_pair, rel=1e-12)


def test_force_error_metric():
    ref = np.array([[3.0, 4.0, 0.0], [0.1, 0.0, 0.0]])
    test = np.array([[3.0, 4.0, 0.5], [0.1, 0.2, 0.0]])
    np.testing.assert_allclose(nonbonded.force_rel_err(test, ref), [0.1, 0.2])


# ------------------------------------------------------------------- bonded


def test_bond_closed_form():
    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.3, 0.4]])
    gx, gp, gl, u = bonded.bonded("bond", x, np.array([[100.0, 0.4]]), 0.0, np.array([[0, 1]]))
    assert u == pytest.approx(100 / 2 * 0.1**2)
    np.testing.assert_allclose(gx[1], 100 * 0.1 * np.array([0, 0.6, 0.8]), rtol=1e-12)
    np.testing.assert_allclose(gp[0], [0.005, -10.0], rtol=1e-12)


def test_angle_closed_form():
    x = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 2.0, 0.0]])  # 90 degrees
    gx, gp, gl, u = bonded.bonded("angle", x, np.array([[50.0, math.pi / 3]]), 0.0, np.array([[0, 1, 2]]))
    assert u == pytest.approx(25.0 * (0.0 - 0.5) ** 2)


def test_torsion_closed_form():
    # trans (phi = pi) and cis (phi = 0) geometries
    trans = np.array([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    cis = np.array([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    ang_t = float(bonded.signed_torsion_angle(*[torch.tensor(v[None]) for v in trans]))
    ang_c = float(bonded.signed_torsion_angle(*[torch.tensor(v[None]) for v in cis]))
    assert abs(abs(ang_t) - math.pi) < 1e-12 and abs(ang_c) < 1e-12
    u = bonded.bonded("torsion", trans, np.array([[3.0, 0.0, 1.0]]), 0.0, np.array([[0, 1, 2, 3]]))[3]
    assert u == pytest.approx(3.0 * (1 + math.cos(math.pi)), abs=1e-12)
    # a 90 degree dihedral has the sign of the reference convention
    q = np.array([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    a = float(bonded.signed_torsion_angle(*[torch.tensor(v[None]) for v in q]))
    assert abs(abs(a) - math.pi / 2) < 1e-12


@pytest.mark.parametrize("kind,arity,params", [("bond", 2, [300.0, 0.12]), ("angle", 3, [200.0, 1.9]), ("torsion", 4, [4.0, 0.7, 2.0])])
def test_bonded_finite_differences(rng, kind, arity, params):
    x = rng.normal(size=(arity, 3)) * 0.2
    idx = np.arange(arity)[None]
    p = np.array([params])
    gx, gp, gl, u = bonded.bonded(kind, x, p, 0.4, idx, mult=[1], offset=[0])
    h = 1e-6
    for i in range(arity):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = (bonded.bonded(kind, xp, p, 0.4, idx, [1], [0])[3] - bonded.bonded(kind, xm, p, 0.4, idx, [1], [0])[3]) / (2 * h)
            assert gx[i, k] == pytest.approx(fd, rel=1e-5, abs=1e-6)
    # du/dlambda = mult * u / prefactor
    assert gl == pytest.approx(u / 0.4, rel=1e-10)


# ------------------------------------------------------------------- neighborlist


def _water(n):
    x = np.load(GOLDEN / "water.npy", allow_pickle=False)[:, :3].astype(np.float32).astype(np.float64)
    x = x[:n]
    box = np.eye(3) * (x.max(0) - x.min(0) + 0.1)
    return x, box


@pytest.mark.parametrize("n", [35, 64, 129, 301])
def test_nblist_reference_equals_brute_force(n):
    x, box = _water(n)
    assert nblist.build_reference_ixn_list(x, box, 1.0, 64) == nblist.brute_force_ixn_list(x, box, 1.0, 64)


@pytest.mark.parametrize("n", [35, 129, 1025, 2029])
def test_nblist_fp32_predicate_matches_f64_away_from_the_boundary(n):
    """The FP32 set differs from the f64 set only on pairs within 1e-5 nm of the cutoff."""
    x, box = _water(n)
    cutoff = 1.0
    ref = nblist.build_reference_ixn_list(x, box, cutoff, 64)
    f32 = nblist.fp32_ixn_list(x, box, cutoff, 64)
    L = np.diag(box)
    for r, (a, b) in enumerate(zip(ref, f32)):
        for j in set(a) ^ set(b):
            rows = x[r * 64 : min(r * 64 + 64, n)]
            d = rows - x[j]
            d -= L * np.floor(d / L + 0.5)
            dmin = np.linalg.norm(d, axis=1).min()
            assert abs(dmin - cutoff) < 1e-5, (r, j, dmin)


def test_nblist_reference_with_32_blocks_matches_reference_definition():
    """Block size 32 reproduces the reference's own definition (tests/test_nblist.py:70-94)."""
    x, box = _water(200)
    lists = nblist.build_reference_ixn_list(x, box, 1.0, 32)
    assert len(lists) == 7
    assert all(all(j >= 32 * r for j in lst) for r, lst in enumerate(lists))
    assert all(set(range(32 * r, min(32 * r + 32, 200))) <= set(lst) for r, lst in enumerate(lists))


# ------------------------------------------------------------------- integrator, barostat


def test_langevin_coefficients_closed_form():
    ca, cb, cc = integrator.langevin_coefficients(300.0, 0.0025, 1.0, np.array([1.0, 16.0]))
    assert ca == pytest.approx(math.exp(-0.0025))
    np.testing.assert_allclose(cb, [0.0025, 0.0025 / 16])
    kT = integrator.BOLTZ * 300
    np.testing.assert_allclose(cc, math.sqrt(1 - math.exp(-0.005)) * np.sqrt(kT / np.array([1.0, 16.0])))


def test_baoab_friction_zero_is_leapfrog():
    """ca = 1, cc = 0: the rotated BAOAB step is leapfrog (v lives at half steps).
    A harmonic oscillator started at rest follows cos(w t) to O(dt) phase and
    its leapfrog energy stays bounded."""
    k, m, dt = 4.0, 1.0, 0.01
    x = np.array([[1.0, 0.0, 0.0]])
    v = np.zeros((1, 3))
    cbs = np.array([-dt / m])
    w = math.sqrt(k / m)
    energies = []
    for n in range(1, 1001):
        x, v = integrator.baoab_step(x, v, k * x, dt, 1.0, cbs, np.zeros(1), np.zeros((1, 3)))
        energies.append(0.5 * k * x[0, 0] ** 2 + 0.5 * m * v[0, 0] ** 2)
        assert x[0, 0] == pytest.approx(math.cos(w * n * dt), abs=2e-2)
    assert max(energies) / min(energies) < 1.03


def test_barostat_scaling_preserves_molecules(rng):
    box = np.eye(3) * 3.0
    groups = [np.arange(3 * i, 3 * i + 3) for i in range(10)]
    x = rng.uniform(0, 3, size=(30, 3))
    x2, box2 = barostat.scale_centroids(x, box, groups, 1.01)
    for g in groups:
        np.testing.assert_allclose(np.diff(x2[g], axis=0), np.diff(x[g], axis=0), atol=1e-12)
        c, c2 = x[g].mean(0), x2[g].mean(0)
        np.testing.assert_allclose(c2 - 1.5, (c - 1.5) * 1.01, atol=1e-12)
    np.testing.assert_allclose(np.diag(box2), 3.03)


def test_barostat_ideal_gas_acceptance():
    """With no energy change the volume move only sees P dV - N kT ln(V'/V)."""
    V, dV = 27.0, 0.2
    w = barostat.acceptance_work(0.0, 0.0, V, dV, 100, 1.0, 300.0)
    p = 1.0 * barostat.AVOGADRO * 1e-25
    assert w == pytest.approx(p * dV - 100 * barostat.BOLTZ * 300 * math.log((V + dV) / V))
    assert barostat.accept(-1.0, 0.999, 300.0)
    assert not barostat.accept(100.0, 0.5, 300.0)
    assert barostat.adapt_volume_scale(1.0, 1, 10, 27.0) == (1.0 / 1.1, 0, 0)
    assert barostat.adapt_volume_scale(1.0, 9, 10, 27.0) == (1.1, 0, 0)


def test_dl_dp_matches_finite_differences_of_du_dl(rng):
    """d(du/dlambda)/dp (double backward) vs central differences of the
    oracle's du/dlambda, which is itself pinned above."""
    x, p, box, lam, excl, sc, pl, of = _tiny_system(rng, N=16)
    g = nonbonded.nonbonded_pairs_dl_dp(x, p, box, lam, excl, sc, 2.0, 1.2, pl, of)
    assert np.abs(g).max() > 0

    def dl(pp):
        return nonbonded.nonbonded_pairs(x, pp, box, lam, excl, sc, 2.0, 1.2, pl, of)[2]

    h = 1e-6
    for i, k in [(0, 0), (1, 1), (2, 2), (4, 0), (7, 1), (9, 2), (12, 0)]:
        pp, pm = p.copy(), p.copy()
        pp[i, k] += h
        pm[i, k] -= h
        assert g[i, k] == pytest.approx((dl(pp) - dl(pm)) / (2 * h), rel=1e-5, abs=1e-6)
    # atoms without an offset and without an offset partner inside the cutoff do not enter
    g0 = nonbonded.nonbonded_pairs_dl_dp(x, p, box, lam, excl, sc, 2.0, 1.2, pl, np.zeros_like(of))
    assert np.all(g0 == 0.0)


def test_bonded_dl_dp_closed_form(rng):
    """With prefactor offset + mult * lambda, d(du/dlambda)/dp = mult * d(shape)/dp."""
    x = rng.normal(size=(6, 3))
    idxs = np.array([[0, 1], [2, 3], [4, 5]])
    params = np.array([[100.0, 0.5], [200.0, 0.3], [50.0, 1.2]])
    mult = np.array([0, 1, -2])
    off = np.array([1, 0, 3])
    g = bonded.bonded_dl_dp("bond", x, params, 0.4, idxs, mult, off)
    r = np.linalg.norm(x[idxs[:, 0]] - x[idxs[:, 1]], axis=1)
    np.testing.assert_allclose(g[:, 0], mult * 0.5 * (r - params[:, 1]) ** 2, rtol=1e-12)
    np.testing.assert_allclose(g[:, 1], -mult * params[:, 0] * (r - params[:, 1]), rtol=1e-12)
    tidx = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    tpar = np.array([[3.0, 0.0, 1.0], [5.0, np.pi, 2.0]])
    gt = bonded.bonded_dl_dp("torsion", x, tpar, 0.7, tidx, np.array([2, 0]), np.array([1, 1]))
    phi = bonded.signed_torsion_angle(*[torch.tensor(x[tidx[:, c]]) for c in range(4)]).numpy()
    np.testing.assert_allclose(gt[0, 0], 2 * (1 + np.cos(tpar[0, 2] * phi[0] - tpar[0, 1])), rtol=1e-12)
    assert np.all(gt[1] == 0.0)


def _interp_params(rng, p0, appear=True):
    """Final parameters for p0; with ``appear`` some LJ-off atoms (eps 0) switch on
    and others off, otherwise the LJ-off atoms stay off."""
    p1 = p0.copy()
    p1[:, 0] = rng.normal(0, 3, p0.shape[0])
    p1[:, 1] = np.clip(p0[:, 1] + rng.normal(0, 0.02, p0.shape[0]), 0.05, 0.2)
    p1[:, 2] = np.abs(p0[:, 2] + rng.normal(0, 0.2, p0.shape[0]))
    p1[p0[:, 2] == 0.0, 2] = 0.0
    if appear:
        p1[::5, 2] = rng.uniform(0.2, 1.0, p1[::5].shape[0])
        p1[::7, 2] = 0.0
    return np.concatenate([p0, p1])


def test_interpolated_endpoints_and_split(rng):
    """At lambda 0 (1) the interpolated energy is the plain energy of the initial
    (final) parameters; du_dp splits as ((1 - lambda) g, lambda g) with g the plain
    du_dp at p(lambda); du_dl gains sum g . (p1 - p0)."""
    x, p0, box, _, excl, sc, pl, of = _tiny_system(rng, N=20)
    P = _interp_params(rng, p0)
    n = len(x)
    for lam, pend in ((0.0, P[:n]), (1.0, P[n:])):
        a = nonbonded.nonbonded_pairs(x, P, box, lam, excl, sc, 2.0, 1.2, pl, of, interpolated=True)
        b = nonbonded.nonbonded_pairs(x, pend, box, lam, excl, sc, 2.0, 1.2, pl, of)
        assert a[3] == pytest.approx(b[3], rel=1e-13)
        np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=1e-12)
    lam = 0.37
    pl_ = (1 - lam) * P[:n] + lam * P[n:]
    gx, gp, gl, u = nonbonded.nonbonded_pairs(x, P, box, lam, excl, sc, 2.0, 1.2, pl, of, interpolated=True)
    hx, hp, hl, v = nonbonded.nonbonded_pairs(x, pl_, box, lam, excl, sc, 2.0, 1.2, pl, of)
    assert gp.shape == (2 * n, 3)
    assert u == pytest.approx(v, rel=1e-13)
    np.testing.assert_allclose(gx, hx, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gp[:n], (1 - lam) * hp, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gp[n:], lam * hp, rtol=1e-12, atol=1e-12)
    assert gl == pytest.approx(hl + np.sum(hp * (P[n:] - P[:n])), rel=1e-11, abs=1e-11)


def test_interpolated_finite_differences(rng):
    """du_dl of the interpolated energy (parameters and the 4D coordinate both move
    with lambda) vs central differences of u in lambda; dense == pair evaluator.
    LJ-off atoms stay off: where eps_i eps_j crosses 0 the LJ term switches (the
    reference masks it, nonbonded.py:227) and differences do not apply."""
    x, p0, box, _, excl, sc, pl, of = _tiny_system(rng, N=16)
    P = _interp_params(rng, p0, appear=False)
    h = 1e-6
    for lam in (0.0, 0.2, 0.7):
        gx, gp, gl, u = nonbonded.nonbonded_v3(x, P, box, lam, excl, sc, 2.0, 1.2, pl, of, interpolated=True)
        b = nonbonded.nonbonded_pairs(x, P, box, lam, excl, sc, 2.0, 1.2, pl, of, interpolated=True)
        assert b[3] == pytest.approx(u, rel=1e-12)
        np.testing.assert_allclose(b[1], gp, rtol=1e-10, atol=1e-10)
        up = nonbonded.nonbonded_v3(x, P, box, lam + h, excl, sc, 2.0, 1.2, pl, of, interpolated=True)[3]
        um = nonbonded.nonbonded_v3(x, P, box, lam - h, excl, sc, 2.0, 1.2, pl, of, interpolated=True)[3]
        assert gl == pytest.approx((up - um) / (2 * h), rel=1e-5, abs=1e-5)
        for i, k in [(0, 0), (3, 1), (17, 2), (30, 0)]:
            pp, pm = P.copy(), P.copy()
            pp[i, k] += h
            pm[i, k] -= h
            fp = nonbonded.nonbonded_v3(x, pp, box, lam, excl, sc, 2.0, 1.2, pl, of, interpolated=True)[3]
            fm = nonbonded.nonbonded_v3(x, pm, box, lam, excl, sc, 2.0, 1.2, pl, of, interpolated=True)[3]
            assert gp[i, k] == pytest.approx((fp - fm) / (2 * h), rel=1e-5, abs=1e-5)


def test_interpolated_equal_endpoints_have_zero_du_dl(rng):
    """Known answer of tests/test_nonbonded.py:303-320: with equal initial and final
    parameters and no lambda offsets, du/dlambda is exactly zero."""
    x, p0, box, lam, excl, sc, pl, of = _tiny_system(rng, N=20, with_offsets=False)
    P = np.concatenate([p0, p0])
    gl = nonbonded.nonbonded_pairs(x, P, box, 0.41, excl, sc, 2.0, 1.2, pl, of, interpolated=True)[2]
    assert gl == 0.0


@pytest.mark.parametrize("interp", [False, True])
def test_interaction_group_is_all_pairs_minus_split_planes(rng, interp):
    """U_AB = U(all) - (U_A + U_B) (tests/nonbonded/test_nonbonded_interaction_group.py:200-260):
    U_A + U_B is the dense all-pairs energy with the row atoms moved to a lambda plane
    farther than the cutoff. Checks u, du_dx, du_dp and du_dl of the pair-restricted
    evaluator against two dense ones."""
    N, L, cutoff = 30, 2.6, 1.1
    x = rng.uniform(0, L, size=(N, 3))
    p0 = np.stack([rng.normal(0, 3, N), rng.uniform(0.08, 0.16, N), rng.uniform(0.1, 1.0, N)], 1)
    P = _interp_params(rng, p0, appear=False) if interp else p0
    box = np.eye(3) * L
    offset = rng.integers(-2, 3, N).astype(np.int32)
    rows = rng.choice(N, 9, replace=False)
    plane0 = np.zeros(N, np.int32)
    plane_split = plane0.copy()
    plane_split[rows] = 5  # |dw| >= (5 - 4) cutoff for every row-column pair
    none = np.zeros((0, 2), int), np.zeros((0, 2))
    for lam in (0.0, 0.1, 0.8):
        ig = nonbonded.interaction_group(x, P, box, lam, rows, 2.0, cutoff, plane0, offset, interpolated=interp)
        a = nonbonded.nonbonded_v3(x, P, box, lam, *none, 2.0, cutoff, plane0, offset, interpolated=interp)
        b = nonbonded.nonbonded_v3(x, P, box, lam, *none, 2.0, cutoff, plane_split, offset, interpolated=interp)
        assert ig[3] != 0.0
        assert ig[3] == pytest.approx(a[3] - b[3], rel=1e-10, abs=1e-10)
        np.testing.assert_allclose(ig[0], a[0] - b[0], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(ig[1], a[1] - b[1], rtol=1e-9, atol=1e-9)
        assert ig[2] == pytest.approx(a[2] - b[2], rel=1e-9, abs=1e-9)


def test_interaction_group_two_particles():
    """One row and one column atom: the pair term; two row atoms: nothing."""
    x = np.array([[0.2, 0.3, 0.4], [0.6, 0.3, 0.4], [0.2, 0.7, 0.4]])
    p = np.array([[1.3, 0.12, 0.8], [-0.7, 0.14, 0.6], [0.5, 0.1, 0.3]])
    box = np.eye(3) * 3.0
    u = nonbonded.interaction_group(x[:2], p[:2], box, 0.0, [0], 2.0, 1.2, [0, 0], [0, 0])[3]
    r, s, e = 0.4, 0.26, 0.48
    assert u == pytest.approx(1.3 * -0.7 * erfc(2 * r) / r + 4 * e * ((s / r) ** 12 - (s / r) ** 6), rel=1e-12)
    # rows {0, 2} and column {1}: the 0-2 pair is not counted
    u3 = nonbonded.interaction_group(x, p, box, 0.0, [0, 2], 2.0, 1.2, [0, 0, 0], [0, 0, 0])[3]
    r12 = math.hypot(0.4, 0.4)
    s12, e12 = 0.24, 0.6 * 0.3
    u12 = -0.7 * 0.5 * erfc(2 * r12) / r12 + 4 * e12 * ((s12 / r12) ** 12 - (s12 / r12) ** 6)
    assert u3 == pytest.approx(u + u12, rel=1e-12)


def test_nblist_subset_reference_equals_brute_force_and_is_order_independent(rng):
    """tests/test_nblist.py:95-193: the row-subset list against a double loop; shuffling
    the row indices changes the blocks but not the union."""
    x = rng.uniform(0, 3.0, size=(300, 3))
    box = np.eye(3) * 3.0
    rows = rng.choice(300, 140, replace=False)
    ref = nblist.build_reference_ixn_list_with_subset(x, box, 0.9, rows, 64)
    assert len(ref) == 3
    L = 3.0
    cols = np.setdiff1d(np.arange(300), rows)
    for rb, lst in enumerate(ref):
        brute = []
        for j in cols:
            d = x[rows[rb * 64 : (rb + 1) * 64]] - x[j]
            d -= L * np.floor(d / L + 0.5)
            if np.any(np.sqrt((d * d).sum(1)) < 0.9):
                brute.append(int(j))
        assert sorted(lst) == brute
    shuffled = rows.copy()
    rng.shuffle(shuffled)
    ref2 = nblist.build_reference_ixn_list_with_subset(x, box, 0.9, shuffled, 64)
    assert set(np.concatenate(ref).tolist()) == set(np.concatenate(ref2).tolist())
    f32 = nblist.fp32_ixn_list_rows(x, box, 0.9, rows, 64)
    assert sum(len(set(a) ^ set(b)) for a, b in zip(ref, f32)) <= 2


# ------------------------------------------------------------- second order


def _grad_all(fn, x, p, lam):
    xt, pt, lt = nonbonded._t(x, True), nonbonded._t(p, True), nonbonded._t(lam, True)
    gx, gp, gl = torch.autograd.grad(fn(xt, pt, lt), (xt, pt, lt), allow_unused=True)
    return gx.numpy(), gp.numpy(), 0.0 if gl is None else float(gl)


def test_hvp_matches_central_differences_of_the_gradient(rng):
    """hvp() (double backward) equals the central difference of the FP64
    gradient along the tangent (vx, vp, vl), for every output: an independent
    route that a wrong sign, dropped term or transposed operand fails."""
    x, params, box, lam, excl, scales, plane, offset = _tiny_system(rng)
    fn = nonbonded.pair_energy_fn(x, box, excl, scales, 2.0, 1.2, plane, offset, pairs=np.array(
        [(i, j) for i in range(len(x)) for j in range(i + 1, len(x))]))
    vx = rng.normal(size=x.shape)
    vp = rng.normal(size=params.shape) * np.array([1.0, 0.01, 0.1])
    # eps == 0 switches LJ off (gradients included, nonbonded.py:214-225): the
    # energy is not smooth in eps there, so the tangent leaves those eps at 0
    vp[params[:, 2] == 0.0, 2] = 0.0
    vl = 0.7
    hx, hp, hl = nonbonded.hvp(fn, x, params, lam, vx, vp, vl)
    h = 1e-6
    gp_ = _grad_all(fn, x + h * vx, params + h * vp, lam + h * vl)
    gm_ = _grad_all(fn, x - h * vx, params - h * vp, lam - h * vl)
    for a, plus, minus in zip((hx, hp, hl), gp_, gm_):
        fd = (np.asarray(plus) - np.asarray(minus)) / (2 * h)
        np.testing.assert_allclose(a, fd, rtol=1e-5, atol=1e-5 * max(np.abs(fd).max(), 1.0))


def test_hvp_is_symmetric(rng):
    """u . H v == v . H u over (x, params, lambda): H is a Hessian."""
    x, params, box, lam, excl, scales, plane, offset = _tiny_system(rng)
    fn = nonbonded.pair_energy_fn(x, box, excl, scales, 2.0, 1.2, plane, offset)
    u = (rng.normal(size=x.shape), rng.normal(size=params.shape), 0.3)
    v = (rng.normal(size=x.shape), rng.normal(size=params.shape), -1.1)
    hu = nonbonded.hvp(fn, x, params, lam, *u)
    hv = nonbonded.hvp(fn, x, params, lam, *v)
    a = sum(float(np.sum(p * q)) for p, q in zip(v, hu))
    b = sum(float(np.sum(p * q)) for p, q in zip(u, hv))
    assert abs(a - b) <= 1e-10 * max(abs(a), 1.0)


def test_bond_hessian_closed_form(rng):
    """U = kb/2 (r - b0)^2: H_ii = kb (n n^T + (1 - b0 / r) (I - n n^T)),
    H_ij = -H_ii; d(du/dx_i)/dkb = (r - b0) n, d(du/dx_i)/db0 = -kb n."""
    x = rng.normal(size=(2, 3))
    kb, b0 = 350.0, 0.11
    d = x[0] - x[1]
    r = np.linalg.norm(d)
    n = d / r
    Hii = kb * (np.outer(n, n) + (1 - b0 / r) * (np.eye(3) - np.outer(n, n)))
    vx = rng.normal(size=(2, 3))
    vp = np.array([[0.3, -0.2]])
    fn = bonded.bonded_energy_fn("bond", np.array([[0, 1]]))
    hx, hp, hl = nonbonded.hvp(fn, x, np.array([[kb, b0]]), 0.0, vx, vp, 0.0)
    want0 = Hii @ (vx[0] - vx[1]) + vp[0, 0] * (r - b0) * n + vp[0, 1] * (-kb * n)
    np.testing.assert_allclose(hx[0], want0, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(hx[1], -want0, rtol=1e-12, atol=1e-12)
    # du/dkb = (r - b0)^2 / 2, du/db0 = -kb (r - b0): their tangents along (vx, vp)
    dr = n @ (vx[0] - vx[1])
    want_p = [(r - b0) * dr - (r - b0) * vp[0, 1], -vp[0, 0] * (r - b0) - kb * (dr - vp[0, 1])]
    np.testing.assert_allclose(hp[0], want_p, rtol=1e-12, atol=1e-12)
    assert hl == 0.0


@pytest.mark.parametrize("kind,arity,params", [
    ("angle", 3, [[75.0, 1.91]]),
    ("torsion", 4, [[2.5, 0.4, 3.0]]),
])
def test_bonded_hvp_central_differences(rng, kind, arity, params):
    """Angle and torsion hvp() (with a lambda-dependent prefactor) against
    central differences of the FP64 gradient."""
    x = rng.normal(size=(arity, 3))
    idxs = np.arange(arity)[None, :]
    fn = bonded.bonded_energy_fn(kind, idxs, mult=[2], offset=[1])
    params = np.asarray(params, dtype=np.float64)
    vx = rng.normal(size=x.shape)
    vp = rng.normal(size=params.shape) * 0.1
    hx, hp, hl = nonbonded.hvp(fn, x, params, 0.4, vx, vp, 0.5)
    h = 1e-6
    gp_ = _grad_all(fn, x + h * vx, params + h * vp, 0.4 + h * 0.5)
    gm_ = _grad_all(fn, x - h * vx, params - h * vp, 0.4 - h * 0.5)
    for a, plus, minus in zip((hx, hp, hl), gp_, gm_):
        fd = (np.asarray(plus) - np.asarray(minus)) / (2 * h)
        np.testing.assert_allclose(a, fd, rtol=1e-5, atol=1e-6 * max(np.abs(fd).max(), 1.0))


def test_trajectory_backprop_matches_finite_differences(rng):
    """oracle.trajectory: the gradient of a loss on the du/dlambda values of a
    short BAOAB trajectory (with noise) equals central differences of the
    forward trajectory in the parameters and in the initial coordinates."""
    from oracle import trajectory

    x, params, box, lam, excl, scales, plane, offset = _tiny_system(rng)
    N = len(x)
    fn = nonbonded.pair_energy_fn(x, box, excl, scales, 2.0, 1.2, plane, offset,
                                  pairs=np.array([(i, j) for i in range(N) for j in range(i + 1, N)]))

    def energy(xt, ps, lt):
        return fn(xt, ps[0], lt)

    masses = rng.uniform(4.0, 16.0, N)
    dt, n_steps, interval = 0.001, 12, 3
    ca, cb, cc = integrator.langevin_coefficients(300.0, dt, 1.0, masses)
    cbs = -cb
    noise = rng.normal(size=(n_steps, N, 3))
    v0 = rng.normal(size=(N, 3)) * 0.5
    dl_bar = rng.normal(size=n_steps // interval)
    dls, (gp,), xb, vb = trajectory.trajectory_loss_grad(energy, x, v0, [params], lam, n_steps, interval, dt, ca,
                                                         cbs, cc, dl_bar, noise)
    assert len(dls) == n_steps // interval

    def loss(p, x0):
        d = trajectory.trajectory_du_dl(energy, x0, v0, [p], lam, n_steps, interval, dt, ca, cbs, cc, noise)
        return float(np.dot(dl_bar, d))

    dp = rng.normal(size=params.shape) * np.array([1.0, 0.01, 0.1])
    dp[params[:, 2] == 0.0, 2] = 0.0
    h = 1e-5
    fd = (loss(params + h * dp, x) - loss(params - h * dp, x)) / (2 * h)
    assert abs(fd - np.sum(gp * dp)) <= 1e-5 * max(abs(fd), 1e-3)
    dx = rng.normal(size=x.shape)
    fdx = (loss(params, x + h * dx) - loss(params, x - h * dx)) / (2 * h)
    assert abs(fdx - np.sum(xb * dx)) <= 1e-5 * max(abs(fdx), 1e-3)


def test_numpy_philox_known_answers_and_normals():
    """oracle.philox: the Random123 known-answer vectors of philox4x32-10, and
    Box-Muller normals with mean 0 and variance 1."""
    from oracle import philox

    kat = [
        ([0, 0, 0, 0], (0, 0), [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, (0xFFFFFFFF, 0xFFFFFFFF), [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], (0xA4093822, 0x299F31D0),
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert philox.philox4x32_10(np.array([ctr], np.uint32), *key)[0].tolist() == want
    z = philox.langevin_noise(1234, 7, 20000)
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.03
