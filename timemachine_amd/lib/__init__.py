"""Picklable integrator and barostat factories with the interface of the
reference's ``timemachine.lib`` (timemachine/lib/__init__.py:7-38)."""

from __future__ import annotations

import numpy as np

from timemachine_amd.lib import custom_ops

# Boltzmann constant in kJ/mol/K as the reference computes it (timemachine/constants.py:3-6).
BOLTZMANN = 1.380658e-23
AVOGADRO = 6.0221367e23
BOLTZ = BOLTZMANN * AVOGADRO / 1000


def langevin_coefficients(temperature, dt, friction, masses):
    """(ca, cb, cc) of the BAOAB integrator (timemachine/integrator.py:10-48):
    ca = exp(-friction dt), cb = dt / m, cc = sqrt(1 - exp(-2 friction dt)) sqrt(kT / m)."""
    masses = np.asarray(masses, dtype=np.float64)
    kT = BOLTZ * temperature
    nscale = np.sqrt(kT / masses)
    ca = np.exp(-friction * dt)
    cb = dt / masses
    cc = np.sqrt(1 - np.exp(-2 * friction * dt)) * nscale
    return ca, cb, cc


class LangevinIntegrator:
    def __init__(self, temperature, dt, friction, masses, seed):
        self.dt = dt
        self.seed = seed
        ca, cb, cc = langevin_coefficients(temperature, dt, friction, masses)
        cb = -cb
        self.ca = ca
        self.cbs = cb
        self.ccs = cc

    def impl(self):
        return custom_ops.LangevinIntegrator(self.dt, self.ca, self.cbs, self.ccs, self.seed)


class MonteCarloBarostat:
    __slots__ = ("N", "temperature", "pressure", "group_idxs", "interval", "seed")

    def __init__(self, N, pressure, temperature, group_idxs, interval, seed):
        self.N = N
        self.pressure = pressure
        self.temperature = temperature
        self.group_idxs = group_idxs
        self.interval = interval
        self.seed = seed

    def impl(self, bound_potentials):
        return custom_ops.MonteCarloBarostat(
            self.N, self.pressure, self.temperature, self.group_idxs, self.interval, bound_potentials, self.seed
        )
