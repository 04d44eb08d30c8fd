"""CPU checks of the window-gradient algebra (no GPU): the estimator from the
observer sums is the derivative of the reweighted window average, and the
trapezoid weights integrate exactly what the trapezoid rule integrates."""

import numpy as np
import pytest

from timemachine_amd.fe import window_gradients as wg


def test_trapezoid_weights():
    lam = np.array([0.0, 0.1, 0.35, 0.7, 1.0])
    f = np.sin(3 * lam) + lam**2
    assert np.dot(wg.trapezoid_weights(lam), f) == pytest.approx(np.trapezoid(f, lam), rel=1e-14)
    np.testing.assert_allclose(wg.trapezoid_weights(np.linspace(0, 1, 5)), [0.125, 0.25, 0.25, 0.25, 0.125])


def test_window_gradient_is_derivative_of_reweighted_average():
    """Synthetic frames with A_f(p) and U_f(p) smooth in p: the estimator
    mean(h) - (mean(A g) - mean(A) mean(g)) / kT equals central differences of
    sum_f A_f(p) w_f(p) / sum_f w_f(p), w_f = exp(-(U_f(p) - U_f(p0)) / kT)."""
    rng = np.random.default_rng(5)
    n_f, P = 40, 6
    kT = 2.49
    a0 = rng.normal(0, 30, n_f)
    ga = rng.normal(0, 5, (n_f, P))  # dA/dp
    u0 = rng.normal(0, 50, n_f)
    gu = rng.normal(0, 20, (n_f, P))  # dU/dp
    cu = rng.normal(0, 3, (n_f, P))  # curvature of U

    def A(p):
        return a0 + ga @ p + 0.5 * np.sin(p).sum()

    def dA(p):
        return ga + 0.5 * np.cos(p)[None, :]

    def U(p):
        return u0 + gu @ p + 0.5 * (cu * p * p).sum(1)

    def dU(p):
        return gu + cu * p[None, :]

    p0 = rng.normal(0, 0.1, P)
    a, g, h = A(p0), dU(p0), dA(p0)
    stats = {"n": n_f, "sum_a": a.sum(), "sum_g": g.sum(0), "sum_ag": (a[:, None] * g).sum(0), "sum_h": h.sum(0)}
    G = wg.window_gradient(stats, kT)

    def rew(p):
        lw = -(U(p) - U(p0)) / kT
        w = np.exp(lw - lw.max())
        return np.sum(A(p) * w) / np.sum(w)

    eps = 1e-5
    for k in range(P):
        e = np.zeros(P)
        e[k] = eps
        fd = (rew(p0 + e) - rew(p0 - e)) / (2 * eps)
        assert G[k] == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_flat_layout_split():
    lay = wg.FlatLayout(n_atoms=4, n_bonds=2, n_angles=1, n_torsions=3)
    flat = np.arange(lay.size, dtype=float)
    parts = lay.split(flat)
    assert parts["nonbonded"].shape == (4, 3) and parts["nonbonded"][3, 2] == 11
    assert parts["bond"].shape == (2, 2) and parts["bond"][0, 0] == 12
    assert parts["angle"].shape == (1, 2) and parts["angle"][0, 1] == 17
    assert parts["torsion"].shape == (3, 3) and parts["torsion"][2, 2] == lay.size - 1


def test_window_gradient_rejects_empty():
    with pytest.raises(RuntimeError):
        wg.window_gradient({"n": 0, "sum_a": 0.0, "sum_g": 0, "sum_ag": 0, "sum_h": 0}, 2.5)
