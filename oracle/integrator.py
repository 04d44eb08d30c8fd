"""Langevin (BAOAB rotated by half a step) reference.

langevin_coefficients restates timemachine/integrator.py:10-48 (with the
Boltzmann constant of timemachine/constants.py:3-6); baoab_step restates
integrator.py:81-88 / cpp/src/integrator.cu:37-69 with cb = -dt/m as
lib/__init__.py:13-17 passes it:
  v_mid = v + cb * du_dx;  v' = ca * v_mid + cc * noise;  x' = x + dt/2 (v_mid + v').

The noise is an input here. Its stream is parity unpinned: the reference draws
it from curand XORWOW (integrator.cu:25-27), a closed library, while the device
draws Philox4x32-10 keyed by (seed, atom, step); only its distribution is
checked (Philox known-answer vectors, Box-Muller moments, friction-0 runs).
"""

from __future__ import annotations

import numpy as np

BOLTZ = 1.380658e-23 * 6.0221367e23 / 1000


def langevin_coefficients(temperature, dt, friction, masses):
    kT = BOLTZ * temperature
    nscale = np.sqrt(kT / np.asarray(masses, dtype=np.float64))
    ca = np.exp(-friction * dt)
    cb = dt / np.asarray(masses, dtype=np.float64)
    cc = np.sqrt(1 - np.exp(-2 * friction * dt)) * nscale
    return ca, cb, cc


def baoab_step(x, v, du_dx, dt, ca, cbs, ccs, noise):
    """One step; cbs are the signed coefficients (-dt/m)."""
    v_mid = v + cbs[:, None] * du_dx
    v_new = ca * v_mid + ccs[:, None] * noise
    x_new = x + 0.5 * dt * (v_mid + v_new)
    return x_new, v_new
