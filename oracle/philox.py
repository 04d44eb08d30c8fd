"""The device's Langevin noise restated in numpy (test infrastructure only).

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11; Random123 constants) over
arrays of counters, and the Box-Muller normals the integrator kernel draws
for atom a of a system at MD step t: counters (2a + 0 | 2a + 1, t low word,
t high word, stream "LANG") under the key (seed low word, seed high word);
the first block gives the x and y normals, the second the z normal. Lets an
FP64 reference follow a GPU trajectory with noise step for step. This is the
device's own stream, not the reference's (curand XORWOW, a closed library):
pinned by the Random123 known-answer vectors and by the device's trajectories.
"""

from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
STREAM_LANGEVIN = 0x4C414E47
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, k0: int, k1: int) -> np.ndarray:
    """ctr [n, 4] uint32 -> [n, 4] uint32."""
    c = np.asarray(ctr, dtype=np.uint64).reshape(-1, 4).copy()
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[:, 0]
        p1 = M1 * c[:, 2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        n = np.empty_like(c)
        n[:, 0] = hi1 ^ c[:, 1] ^ np.uint64(k0)
        n[:, 1] = lo1
        n[:, 2] = hi0 ^ c[:, 3] ^ np.uint64(k1)
        n[:, 3] = lo0
        c = n
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c.astype(np.uint32)


def _u01_open0(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    k = ((a.astype(np.uint64) >> np.uint64(5)) << np.uint64(26)) | (b.astype(np.uint64) >> np.uint64(6))
    return (k.astype(np.float64) + 1.0) * 2.0**-53


def normal2(r: np.ndarray):
    u1 = _u01_open0(r[:, 0], r[:, 1])
    u2 = _u01_open0(r[:, 2], r[:, 3])
    rad = np.sqrt(-2.0 * np.log(u1))
    ang = 6.283185307179586476925286766559 * u2
    return rad * np.cos(ang), rad * np.sin(ang)


def langevin_noise(seed: int, step: int, n_atoms: int) -> np.ndarray:
    """[n_atoms, 3] standard normals of MD step `step` of a system with this seed."""
    a = np.arange(n_atoms, dtype=np.uint64)
    lo, hi = int(step) & 0xFFFFFFFF, (int(step) >> 32) & 0xFFFFFFFF
    k0, k1 = int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF
    out = np.empty((n_atoms, 3))
    for blk in range(2):
        ctr = np.stack([2 * a + np.uint64(blk), np.full(n_atoms, lo, np.uint64), np.full(n_atoms, hi, np.uint64),
                        np.full(n_atoms, STREAM_LANGEVIN, np.uint64)], 1)
        z0, z1 = normal2(philox4x32_10(ctr, k0, k1))
        if blk == 0:
            out[:, 0], out[:, 1] = z0, z1
        else:
            out[:, 2] = z0
    return out
