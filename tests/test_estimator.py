"""BAR / MBAR free-energy estimates (CPU): the reference's fixtures for its
helpers (tests/test_fe_utils.py:9-38), and the estimators against closed
forms: Gaussian work distributions, harmonic oscillators, and MBAR reducing to
BAR for two states."""

import math

import numpy as np
import pytest

from timemachine_amd.fe import estimator as est


def test_sanitize_energies_reference_fixture():
    full_us = np.array([[15000.0, -5081923.0, 1598, 1.5, -23.0], [-423581.0, np.nan, -238, 13.5, 23.0]])
    expected = np.array([[np.inf, np.inf, 1598, 1.5, -23.0], [np.inf, np.inf, -238, 13.5, 23.0]])
    np.testing.assert_array_equal(est.sanitize_energies(full_us, lamb_idx=3, cutoff=10000), expected)


def test_extract_delta_Us():
    rng = np.random.default_rng(0)
    U = rng.random((4, 8, 4))
    d = est.extract_delta_Us_from_U_knk(U)
    exp = np.array([[U[i, :, i + 1] - U[i, :, i], U[i + 1, :, i] - U[i + 1, :, i + 1]] for i in range(3)])
    np.testing.assert_array_equal(d, exp)


def _gauss_works(rng, dF, sigma, nF, nR):
    """Forward / reverse works consistent with the Crooks relation for a Gaussian
    forward distribution: w_F ~ N(dF + s^2/2, s^2), w_R ~ N(-dF + s^2/2, s^2)."""
    return rng.normal(dF + sigma**2 / 2, sigma, nF), rng.normal(-dF + sigma**2 / 2, sigma, nR)


def test_bar_gaussian_works_and_error_estimate():
    rng = np.random.default_rng(1)
    dF, sigma = 3.7, 1.5
    ests, errs = [], []
    for _ in range(60):
        wF, wR = _gauss_works(rng, dF, sigma, 800, 500)
        d, e = est.bar(wF, wR)
        ests.append(d)
        errs.append(e)
    ests = np.array(ests)
    assert abs(ests.mean() - dF) < 4 * ests.std() / math.sqrt(len(ests))
    # the asymptotic error matches the spread of the estimates
    assert np.mean(errs) == pytest.approx(ests.std(), rel=0.25)


def test_bar_is_exact_for_symmetric_identical_states():
    w = np.random.default_rng(2).normal(0.0, 0.7, 400)
    d, _ = est.bar(w, w)
    assert abs(d) < 1e-10


def _oscillators(rng, K_k, O_k, N_k):
    """Samples of 1D harmonic oscillators u_k(x) = K_k/2 (x - O_k)^2 (kT units),
    reduced energies of every sample in every state, and the exact f_k."""
    x = np.concatenate([rng.normal(o, 1 / math.sqrt(k), n) for k, o, n in zip(K_k, O_k, N_k)])
    u_kn = 0.5 * np.asarray(K_k)[:, None] * (x[None, :] - np.asarray(O_k)[:, None]) ** 2
    f = -np.log(np.sqrt(2 * math.pi / np.asarray(K_k)))
    return u_kn, f - f[0]


def test_mbar_harmonic_oscillators():
    rng = np.random.default_rng(3)
    K_k = [16.0, 9.0, 4.0, 1.0]
    O_k = [0.0, 0.3, 0.5, 1.0]
    N_k = [1500, 1500, 1500, 1500]
    est_f, err_f = [], []
    for _ in range(25):
        u_kn, f_exact = _oscillators(rng, K_k, O_k, N_k)
        m = est.MBAR(u_kn, N_k)
        D, E = m.getFreeEnergyDifferences()
        est_f.append(D[0])
        err_f.append(E[0])
    est_f, err_f = np.array(est_f), np.array(err_f)
    for j in range(1, 4):
        assert abs(est_f[:, j].mean() - f_exact[j]) < 4 * est_f[:, j].std() / math.sqrt(len(est_f))
        assert err_f[:, j].mean() == pytest.approx(est_f[:, j].std(), rel=0.35)


def test_mbar_two_states_equals_bar():
    rng = np.random.default_rng(4)
    u_kn, _ = _oscillators(rng, [4.0, 2.0], [0.0, 0.4], [700, 900])
    m = est.MBAR(u_kn, [700, 900])
    D, E = m.getFreeEnergyDifferences()
    wF = u_kn[1, :700] - u_kn[0, :700]
    wR = u_kn[0, 700:] - u_kn[1, 700:]
    d, e = est.bar(wF, wR)
    assert D[0, 1] == pytest.approx(d, abs=1e-8)
    assert E[0, 1] == pytest.approx(e, rel=1e-6)


def test_deltaG_pipeline_on_oscillator_windows():
    """[K, F, K] energies of K oscillator windows (kJ/mol at beta = 1 / 2.494):
    BAR over neighbours and MBAR both recover the exact free energy."""
    rng = np.random.default_rng(5)
    beta = 1 / 2.494
    K_k = np.linspace(9.0, 2.0, 6)
    O_k = np.linspace(0.0, 0.8, 6)
    F = 2000
    x = np.stack([rng.normal(o, 1 / math.sqrt(k), F) for k, o in zip(K_k, O_k)])  # [K, F]
    U = 0.5 * K_k[None, None, :] * (x[:, :, None] - O_k[None, None, :]) ** 2 / beta  # [K, F, K]
    f = -np.log(np.sqrt(2 * math.pi / K_k))
    exact = (f[-1] - f[0]) / beta
    r = est.deltaG_from_U_knk(U, beta, np.linspace(0, 1, 6))
    assert len(r["pairs"]) == 5
    assert abs(r["bar_dG"] - exact) < 4 * r["bar_dG_err"]
    assert abs(r["mbar_dG"] - exact) < 4 * r["mbar_dG_err"]
