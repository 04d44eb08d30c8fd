"""Free-energy estimates from lambda-window energies: BAR and MBAR.

The reference turns the [K, F, K] energies of Context.multiple_steps_U
(every window's frames evaluated at every window) into a free energy in
fe/estimator.py:284-423 (deltaG_from_results), with the helpers of
fe/utils.py:284-364 and the BAR / MBAR estimators of pymbar 3.0.5
(requirements.txt:5), which is not installed here. The estimators below are
written from their published definitions:

* BAR: Bennett, J. Comput. Phys. 22, 245 (1976); Shirts, Bair, Hooker,
  Pande, PRL 91, 140601 (2003). With forward works w_F (n_F samples) and
  reverse works w_R (n_R samples) in kT and M = ln(n_F / n_R), dF solves
      sum_F 1 / (1 + exp(M + w_F - dF)) = sum_R 1 / (1 + exp(-M + w_R + dF)),
  and its asymptotic variance is
      (<f_F^2> / <f_F>^2 - 1) / n_F + (<f_R^2> / <f_R>^2 - 1) / n_R
  with f_F = 1 / (1 + exp(M + w_F - dF)), f_R = 1 / (1 + exp(-M + w_R + dF)).
* MBAR: Shirts, Chodera, J. Chem. Phys. 129, 124105 (2008). Dimensionless
  free energies f_i = -ln sum_n exp(-u_i(x_n)) / sum_k N_k exp(f_k - u_k(x_n))
  with f_0 = 0, solved here by self-consistent iteration followed by Newton
  steps; the covariance of f is W^T (I - W N W^T)^+ W (eq. D8) with
  W_nk = exp(f_k - u_k(x_n)) / sum_j N_j exp(f_j - u_j(x_n)), evaluated
  through the thin SVD of W so that only K x K matrices are inverted.

All arithmetic is FP64 NumPy on the host: the inputs are kilobytes per
window, produced by the device run.
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
from scipy.optimize import brentq
from scipy.special import logsumexp


def sanitize_energies(full_us: np.ndarray, lamb_idx: int, cutoff: float = 10000.0) -> np.ndarray:
    """fe/utils.py:284-321: entries of the [F, K] matrix that differ from column
    lamb_idx by cutoff or more (or are NaN) become +inf."""
    full_us = np.asarray(full_us, dtype=np.float64)
    ref = full_us[:, lamb_idx][:, None]
    keep = np.abs(full_us - ref) < cutoff
    return np.where(keep, full_us, np.inf)


def extract_delta_Us_from_U_knk(U_knk: np.ndarray) -> np.ndarray:
    """fe/utils.py:324-364: [K - 1, 2, F] forward (window i's frames at i + 1
    minus at i) and reverse (window i + 1's frames at i minus at i + 1) energy
    differences of a [K, F, K] matrix."""
    U_knk = np.asarray(U_knk, dtype=np.float64)
    K = U_knk.shape[0]
    if U_knk.shape[-1] != K:
        raise ValueError("U_knk must be [K, F, K]")
    out = np.empty((K - 1, 2, U_knk.shape[1]))
    for i in range(K - 1):
        out[i, 0] = U_knk[i, :, i + 1] - U_knk[i, :, i]
        out[i, 1] = U_knk[i + 1, :, i] - U_knk[i + 1, :, i + 1]
    return out


def _log_fermi(z: np.ndarray) -> np.ndarray:
    """log(1 / (1 + exp(z))) without overflow."""
    return -np.logaddexp(0.0, z)


def _bar_zero(w_F: np.ndarray, w_R: np.ndarray, dF: float, M: float) -> float:
    return float(logsumexp(_log_fermi(M + w_F - dF)) - logsumexp(_log_fermi(-M + w_R + dF)))


def bar(w_F: Sequence[float], w_R: Sequence[float], compute_uncertainty: bool = True,
        relative_tolerance: float = 1e-12) -> Tuple[float, Optional[float]]:
    """(dF, d dF) in kT from forward and reverse works in kT (the values
    pymbar.BAR(w_F, w_R) returns in the reference's call, fe/estimator.py:344)."""
    w_F = np.asarray(w_F, dtype=np.float64).reshape(-1)
    w_R = np.asarray(w_R, dtype=np.float64).reshape(-1)
    n_F, n_R = w_F.size, w_R.size
    if n_F == 0 or n_R == 0:
        raise ValueError("BAR needs forward and reverse samples")
    M = math.log(n_F / n_R)
    # bracket: the two exponential-averaging estimates, widened until the sign changes
    lo = float(logsumexp(-w_R) - math.log(n_R))  # dF from the reverse works
    hi = float(-(logsumexp(-w_F) - math.log(n_F)))  # dF from the forward works
    lo, hi = min(lo, hi), max(lo, hi)
    span = max(hi - lo, 1.0)
    g = lambda d: _bar_zero(w_F, w_R, d, M)  # noqa: E731  increasing in d
    while g(lo) > 0:
        lo -= span
        span *= 2
    while g(hi) < 0:
        hi += span
        span *= 2
    dF = brentq(g, lo, hi, xtol=1e-14, rtol=max(relative_tolerance, 4 * np.finfo(float).eps), maxiter=500)
    if not compute_uncertainty:
        return dF, None
    lf = _log_fermi(M + w_F - dF)
    lr = _log_fermi(-M + w_R + dF)
    ratio_F = math.exp(logsumexp(2 * lf) - math.log(n_F) - 2 * (logsumexp(lf) - math.log(n_F)))
    ratio_R = math.exp(logsumexp(2 * lr) - math.log(n_R) - 2 * (logsumexp(lr) - math.log(n_R)))
    var = (ratio_F - 1.0) / n_F + (ratio_R - 1.0) / n_R
    return dF, math.sqrt(max(var, 0.0))


class MBAR:
    """MBAR over reduced energies u_kn [K, N] (state k evaluated on every sample
    n, samples ordered by the state that drew them) and counts N_k."""

    def __init__(self, u_kn: np.ndarray, N_k: Sequence[int], relative_tolerance: float = 1e-12,
                 maximum_iterations: int = 10000):
        self.u_kn = np.asarray(u_kn, dtype=np.float64)
        self.N_k = np.asarray(N_k, dtype=np.float64)
        K, N = self.u_kn.shape
        if self.N_k.size != K or int(self.N_k.sum()) != N:
            raise ValueError("N_k must have one count per state and sum to the number of samples")
        logN = np.where(self.N_k > 0, np.log(np.maximum(self.N_k, 1e-300)), -np.inf)
        f = np.zeros(K)
        for it in range(maximum_iterations):
            log_den = logsumexp(f[:, None] - self.u_kn + logN[:, None], axis=0)  # [N]
            f_new = -logsumexp(-self.u_kn - log_den[None, :], axis=1)
            f_new -= f_new[0]
            if it > 20:
                # then Newton steps on the convex MBAR objective, whose stationary
                # point is the solution of the same equations
                f_new = self._newton(f_new, logN)
            delta = np.max(np.abs(f_new - f))
            f = f_new
            if delta <= relative_tolerance * max(1.0, np.max(np.abs(f))):
                break
        self.f_k = f
        self._logN = logN

    def _log_weights(self, f: np.ndarray, logN: np.ndarray) -> np.ndarray:
        log_den = logsumexp(f[:, None] - self.u_kn + logN[:, None], axis=0)
        return f[:, None] - self.u_kn - log_den[None, :]  # log W^T [K, N]

    def _newton(self, f: np.ndarray, logN: np.ndarray) -> np.ndarray:
        # objective sum_n ln sum_k N_k exp(f_k - u_kn) - sum_k N_k f_k:
        # gradient N_k (sum_n W_kn - 1), Hessian diag(N_k sum_n W_kn) - (N W)(N W)^T
        W = np.exp(self._log_weights(f, logN))  # [K, N]
        grad = self.N_k * W.sum(axis=1) - self.N_k
        H = np.diag(self.N_k * W.sum(axis=1)) - (self.N_k[:, None] * W) @ (self.N_k[:, None] * W).T
        # f_0 fixed at zero: solve the reduced system
        step = np.zeros_like(f)
        step[1:] = np.linalg.solve(H[1:, 1:], -grad[1:])
        out = f + step
        return out - out[0]

    def covariance(self) -> np.ndarray:
        """Theta = W^T (I - W N W^T)^+ W, through the thin SVD of W [N, K]."""
        W = np.exp(self._log_weights(self.f_k, self._logN)).T  # [N, K]
        U, s, Vt = np.linalg.svd(W, full_matrices=False)
        Sig = np.diag(s)
        A = Sig @ Vt @ np.diag(self.N_k) @ Vt.T @ Sig
        inner = np.linalg.pinv(np.eye(len(s)) - A, rcond=1e-10)
        return Vt.T @ Sig @ inner @ Sig @ Vt

    def getFreeEnergyDifferences(self) -> Tuple[np.ndarray, np.ndarray]:
        """(Delta f_ij = f_j - f_i, its standard error) like pymbar's MBAR."""
        f = self.f_k
        D = f[None, :] - f[:, None]
        T = self.covariance()
        d = np.diag(T)
        var = d[:, None] + d[None, :] - 2 * T
        return D, np.sqrt(np.maximum(var, 0.0))


def deltaG_from_U_knk(U_knk: np.ndarray, beta: float, lambda_schedule: Optional[Sequence[float]] = None,
                      cutoff: float = 10000.0) -> Dict[str, object]:
    """The estimates of deltaG_from_results (fe/estimator.py:326-401, without the
    endpoint correction) from a [K, F, K] energy matrix in kJ/mol: BAR summed
    over adjacent windows with errors added in quadrature, and MBAR over all
    windows on sanitised energies. Energies are multiplied by beta (1/kT)."""
    U_knk = np.asarray(U_knk, dtype=np.float64)
    K, F, _ = U_knk.shape
    dU = extract_delta_Us_from_U_knk(U_knk)
    pairs = []
    dG = 0.0
    var = 0.0
    for i in range(K - 1):
        d, e = bar(beta * dU[i, 0], beta * dU[i, 1])
        pairs.append({"from": None if lambda_schedule is None else float(lambda_schedule[i]),
                      "to": None if lambda_schedule is None else float(lambda_schedule[i + 1]),
                      "dG": d / beta, "dG_err": e / beta})
        dG += d / beta
        var += (e / beta) ** 2
    clean = np.array([sanitize_energies(U_knk[k], k, cutoff) for k in range(K)])  # [K, F, K]
    u_kn = beta * clean.reshape(-1, K).T  # [K, K * F]
    mbar = MBAR(u_kn, [F] * K)
    D, E = mbar.getFreeEnergyDifferences()
    return {"bar_dG": dG, "bar_dG_err": math.sqrt(var), "mbar_dG": float(D[0, -1] / beta),
            "mbar_dG_err": float(E[0, -1] / beta), "pairs": pairs}


def simulate_windows(system, lambda_schedule: Sequence[float], equil_steps: int, prod_steps: int,
                     u_interval: int, seed: int = 2022, temperature: float = 300.0, dt: float = 0.0025,
                     friction: float = 1.0, barostat_interval: int = 25, minimize_steps: int = 99,
                     x_interval: int = 0):
    """The per-window protocol of the reference's simulate (fe/estimator.py:200-282)
    for every window of the schedule at once, in one batch on the device: FIRE
    minimisation at the window's lambda and the force check, equilibration, then
    production with the energies of every stored frame at every window
    (multiple_steps_U). Returns (U_knk [K, F, K] in kJ/mol, frames, boxes)."""
    from timemachine_amd import testsystems as ts
    from timemachine_amd.lib import LangevinIntegrator, MonteCarloBarostat, custom_ops
    from timemachine_amd.md.minimizer import FORCE_LIMIT

    lams = np.asarray(lambda_schedule, dtype=np.float64).reshape(-1)
    bps = [p.bound_impl(np.float32) for p in system.bound_potentials()]
    ctxs, keep = [], []
    for k in range(lams.size):
        s = seed * 1009 + k
        intg = LangevinIntegrator(temperature, dt, friction, system.masses, s).impl()
        baro = None
        if barostat_interval > 0:
            baro = MonteCarloBarostat(system.N, 1.0, temperature, system.groups, barostat_interval, s + 17).impl(bps)
        v0 = ts.maxwell_boltzmann(system.masses, temperature, s)
        ctxs.append(custom_ops.Context(system.x, v0, system.box, intg, bps, barostat=baro))
        keep.append((intg, baro))
    batch = custom_ops.BatchedContexts(ctxs, lams)
    if minimize_steps > 0:
        batch.fire_minimize(n_steps=minimize_steps)
        if not np.all(batch.max_force() < FORCE_LIMIT):
            raise AssertionError("Forces much greater than expected after minimization")
    if equil_steps > 0:
        batch.multiple_steps(equil_steps)
    u, frames, boxes = batch.multiple_steps_U(prod_steps, lams, u_interval, x_interval)
    del batch, keep
    return u, frames, boxes
