"""Lambda-window batches of one system: W windows x R replicas in one engine.

The free-energy layer of the reference runs one Context per lambda window,
one process per window (fe/estimator.py:172-281, parallel/client.py:122-153).
Here every window and replica of a system is a system of one batched engine,
built in one call (BatchedContexts.from_windows) instead of a Context per
window. Window k (replica k // W, lambda lambdas[k % W]) is seeded so that the
same window can also be built alone through the Context API
(``window_context``) and stepped to the same bits.
"""

from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from timemachine_amd import testsystems as ts
from timemachine_amd.lib import LangevinIntegrator, MonteCarloBarostat, custom_ops

TEMPERATURE = 300.0
DT = 0.0025
FRICTION = 1.0
PRESSURE_BAR = 1.0


def window_layout(lambdas: Sequence[float], replicas: int, seed0: int = 2022) -> Tuple[List[float], List[int]]:
    """Lambda and integrator seed of every system, replica-major (the
    barostat seed is the integrator seed + 7, the velocities are drawn from
    the integrator seed)."""
    lams, seeds = [], []
    for r in range(replicas):
        for lam in lambdas:
            seeds.append(seed0 + 1000 * r + len(seeds))
            lams.append(float(lam))
    return lams, seeds


def build_batch(sys_: ts.System, lambdas: Sequence[float], replicas: int, seed0: int = 2022,
                barostat_interval: int = 15, minimize_steps: int = 99) -> custom_ops.BatchedContexts:
    """All windows x replicas of ``sys_`` in one engine (BAOAB at 2.5 fs, 300 K,
    1/ps; MC barostat at 1 bar every ``barostat_interval`` steps, 0 = none).
    Every window is first FIRE-minimised at its own lambda for
    ``minimize_steps`` steps on the device, as the reference's simulate does
    before its MD (fe/estimator.py:242-249): the synthetic boxes start with
    close contacts that 2.5 fs steps do not survive. Minimisation leaves
    velocities, box, step counters and barostat state alone."""
    lams, seeds = window_layout(lambdas, replicas, seed0)
    bps = [p.bound_impl(np.float32) for p in sys_.bound_potentials()]
    intg = LangevinIntegrator(TEMPERATURE, DT, FRICTION, sys_.masses, seeds[0]).impl()
    baro = None
    if barostat_interval > 0:
        baro = MonteCarloBarostat(sys_.N, PRESSURE_BAR, TEMPERATURE, sys_.groups, barostat_interval,
                                  seeds[0] + 7).impl(bps)
    v0s = np.stack([ts.maxwell_boltzmann(sys_.masses, TEMPERATURE, sd) for sd in seeds])
    batch = custom_ops.BatchedContexts.from_windows(sys_.x, v0s, sys_.box, intg, seeds, bps, lams, barostat=baro,
                                                    barostat_seeds=[sd + 7 for sd in seeds])
    if minimize_steps > 0:
        batch.fire_minimize(n_steps=minimize_steps)
    return batch


def window_context(sys_: ts.System, seed: int, barostat_interval: int = 15, x0=None):
    """One window alone through the Context API, seeded as ``build_batch`` seeds
    the window with integrator seed ``seed``, starting from ``x0`` (default: the
    system's coordinates; pass a minimised batch window's coordinates to
    continue it alone). Returns (context, handles to keep)."""
    bps = [p.bound_impl(np.float32) for p in sys_.bound_potentials()]
    intg = LangevinIntegrator(TEMPERATURE, DT, FRICTION, sys_.masses, seed).impl()
    baro = None
    if barostat_interval > 0:
        baro = MonteCarloBarostat(sys_.N, PRESSURE_BAR, TEMPERATURE, sys_.groups, barostat_interval,
                                  seed + 7).impl(bps)
    v0 = ts.maxwell_boltzmann(sys_.masses, TEMPERATURE, seed)
    ctx = custom_ops.Context(sys_.x if x0 is None else x0, v0, sys_.box, intg, bps, barostat=baro)
    return ctx, (bps, intg, baro)


def ti_free_energy(lambdas: Sequence[float], du_dl: np.ndarray, replicas: int) -> Tuple[float, np.ndarray]:
    """TI estimate (trapezoid over the windows, fe/math_utils.py:4) from the
    du/dlambda series [W * replicas, K] of a replica-major batch: the per-window
    mean over replicas and frames, then the trapezoid. Returns (dG, per-window means)."""
    lam = np.asarray(lambdas, dtype=np.float64)
    W = lam.size
    means = np.asarray(du_dl, dtype=np.float64).reshape(replicas, W, -1).mean(axis=(0, 2))
    return float(np.sum(0.5 * (means[1:] + means[:-1]) * np.diff(lam))), means
