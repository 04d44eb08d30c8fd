"""FIRE minimisation (test infrastructure only).

Restates the FIRE descent of timemachine/_vendored/fire.py:60-140 (from
jax-md; Bitzek et al., PRL 97, 170201 (2006)) as the reference's
md/minimizer.py:35-73 drives it, with the reference's numpy types: the
constants pass through np.min_scalar_type (all become float16), so dt and
alpha are float16 and update by float16 products, dt**2 is a float16 square
(0 while dt < 1.7e-4), dt * float32(0.5) is float32, and positions,
velocities and forces are float64.
"""

from __future__ import annotations

from typing import Callable, List, Optional

import numpy as np

F16 = np.float16
DT_START = F16(1e-5)
DT_MAX = F16(1e-3)
N_MIN = 5
F_INC = F16(1.1)
F_DEC = F16(0.5)
ALPHA_START = F16(0.1)
F_ALPHA = F16(0.99)
F_EPS = np.float64(np.float32(1e-6))


def fire_minimize(x0: np.ndarray, force: Callable[[np.ndarray, float], np.ndarray], lamb_sched,
                  trace: Optional[List[dict]] = None) -> np.ndarray:
    """Positions after len(lamb_sched) - 1 steps; force(x, lamb) returns -dU/dx.
    ``trace`` (if given) receives dt, alpha, n_pos and P after every step."""
    lamb_sched = np.asarray(lamb_sched, dtype=np.float64).reshape(-1)
    R = np.array(x0, dtype=np.float64)
    V = np.zeros_like(R)
    F = np.asarray(force(R, lamb_sched[0]), dtype=np.float64)
    dt, alpha, n_pos = DT_START, ALPHA_START, 0
    for lamb in lamb_sched[1:]:
        dt_sq = np.square(dt)  # float16
        R = R + (np.float64(dt) * V + np.float64(dt_sq) * F)
        F_new = np.asarray(force(R, lamb), dtype=np.float64)
        V = V + np.float64(np.float32(dt) * np.float32(0.5)) * (F + F_new)
        F = F_new
        f_norm = np.sqrt(np.sum(F * F) + F_EPS)
        v_norm = np.sqrt(np.sum(V * V))
        P = float(np.dot(F.reshape(-1), V.reshape(-1)))
        V = V + np.float64(alpha) * (F * v_norm / f_norm - V)
        n_pos = n_pos + 1 if P >= 0 else 0
        if P > 0 and n_pos > N_MIN:
            dt = min(F16(dt * F_INC), DT_MAX)
            alpha = F16(alpha * F_ALPHA)
        if P < 0:
            dt = F16(dt * F_DEC)
            alpha = ALPHA_START
            V = np.zeros_like(V)
        if trace is not None:
            trace.append({"dt": dt, "alpha": alpha, "n_pos": n_pos, "P": P})
    return R
