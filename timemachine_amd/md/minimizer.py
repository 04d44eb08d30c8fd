"""FIRE minimisation on the device, with the interface of the reference's
timemachine/md/minimizer.py:35-163.

``fire_minimize`` keeps the reference's signature and numerics (the FIRE
variant of _vendored/fire.py:60-140, including its binary16 step size and
mixing factor) but runs every step on the GPU: the force, the FIRE update and
the neighborlist stay in HBM and nothing returns to the host until the end.
``BatchedContexts.fire_minimize`` does the same for all systems of a batch at
once (each with its own FIRE state).

``minimize_4d`` is the device part of minimize_host_4d (md/minimizer.py:76-163)
once the host and ligand potentials are bound: FIRE at lambda = 1, 0 K
Langevin dynamics from lambda 1 to 0 with the fixed atoms made 1e5 times
heavier, FIRE at lambda = 0, and the force-norm check. Building the system
from RDKit molecules and an OpenMM host is not part of this package.
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from timemachine_amd import _lib
from timemachine_amd._lib import call, pd
from timemachine_amd.lib import LangevinIntegrator, custom_ops

FORCE_LIMIT = 25000.0  # kJ/mol/nm, md/minimizer.py:158-161 and fe/estimator.py:245-249


def fire_minimize(x0: np.ndarray, u_impls: Sequence[custom_ops.BoundPotential], box: np.ndarray,
                  lamb_sched: np.ndarray) -> np.ndarray:
    """Minimised coordinates after len(lamb_sched) - 1 FIRE steps, the k-th at
    lamb_sched[k] (md/minimizer.py:35-73)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    sched = np.ascontiguousarray(lamb_sched, dtype=np.float64).reshape(-1)
    N = x0.shape[0]
    out = np.zeros_like(x0)
    arr = _lib.handle_array([bp.handle for bp in u_impls])
    call("tm_fire_minimize", N, pd(x0), pd(box), len(u_impls), arr, sched.size, pd(sched), pd(out))
    return out


def check_forces(x: np.ndarray, u_impls: Sequence[custom_ops.BoundPotential], box: np.ndarray, lamb: float,
                 limit: float = FORCE_LIMIT) -> float:
    """Raise if any potential's per-atom force norm reaches ``limit``
    (fe/estimator.py:245-249); returns the largest norm."""
    worst = 0.0
    for bp in u_impls:
        du_dx, _, _ = bp.execute(x, box, lamb)
        worst = max(worst, float(np.linalg.norm(du_dx, axis=-1).max()))
    if not worst < limit:
        raise AssertionError("Forces much greater than expected after minimization")
    return worst


def minimize_4d(x0: np.ndarray, u_impls: Sequence[custom_ops.BoundPotential], masses: np.ndarray,
                box: np.ndarray, fixed_idxs: Optional[np.ndarray] = None, n_fire: int = 50, n_anneal: int = 1000,
                dt: float = 1.5e-3, seed: int = 0) -> np.ndarray:
    """minimize_host_4d (md/minimizer.py:76-163) on bound potentials: returns all
    coordinates (the caller keeps the host part)."""
    masses = np.asarray(masses, dtype=np.float64).copy()
    if fixed_idxs is not None:
        masses[np.asarray(fixed_idxs)] *= 100000  # keeps the ligand in place
    x = fire_minimize(x0, u_impls, box, np.ones(n_fire))
    intg = LangevinIntegrator(0.0, dt, 1.0, masses, seed).impl()
    ctx = custom_ops.Context(x, np.zeros_like(x), box, intg, list(u_impls))
    _, xs, _ = ctx.multiple_steps(np.linspace(1.0, 0, n_anneal))
    x = fire_minimize(xs[-1], u_impls, box, np.zeros(n_fire))
    check_forces(x, u_impls, box, 0.0)
    return x


__all__ = ["fire_minimize", "check_forces", "minimize_4d", "FORCE_LIMIT"]
