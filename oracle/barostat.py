"""Monte Carlo barostat reference (timemachine/cpp/src/barostat.cu:139-373).

Pressure in internal units is bar * AVOGADRO * 1e-25 (barostat.cu:350) and
kT uses BOLTZ = 0.008314462618 (barostat.cu:12). The volume proposals and
acceptance draws are inputs here; their stream is parity unpinned (curand in
the reference, barostat.cu:298; Philox4x32-10 on the device), only the
volume distribution is checked.
"""

from __future__ import annotations

import numpy as np

AVOGADRO = 6.0221367e23
BOLTZ = 0.008314462618


def scale_centroids(x, box, groups, length_scale):
    """rescale_positions (barostat.cu:139-179): each group's centroid c moves to
    (c - centre) * s + centre, the group rigidly with it; box diagonal * s."""
    x = np.array(x, dtype=np.float64)
    centre = np.diag(box) * 0.5
    for g in groups:
        g = np.asarray(g)
        c = x[g].mean(0)
        x[g] += (c - centre) * length_scale + centre - c
    new_box = np.array(box, dtype=np.float64)
    new_box[np.diag_indices(3)] *= length_scale
    return x, new_box


def acceptance_work(u_before, u_after, volume, delta_volume, n_molecules, pressure_bar, temperature):
    """w = dU + P dV - N_mol kT ln(V'/V) (barostat.cu:241-244)."""
    p = pressure_bar * AVOGADRO * 1e-25
    kT = BOLTZ * temperature
    return (u_after - u_before) + p * delta_volume - n_molecules * kT * np.log((volume + delta_volume) / volume)


def accept(w, r2, temperature):
    """Metropolis: reject iff w > 0 and r2 > exp(-w / kT) (barostat.cu:246)."""
    kT = BOLTZ * temperature
    return not (w > 0 and r2 > np.exp(-w / kT))


def adapt_volume_scale(vscale, accepted, attempted, volume):
    """Step-size adaptation after each attempt (barostat.cu:251-265)."""
    if attempted >= 10:
        if accepted < 0.25 * attempted:
            return vscale / 1.1, 0, 0
        if accepted > 0.75 * attempted:
            return min(vscale * 1.1, volume * 0.3), 0, 0
    return vscale, accepted, attempted
