"""FP64 bonded terms (Torch, autograd derivatives).

Restates timemachine/potentials/bonded.py: harmonic_bond (:27-80),
harmonic_angle with cos_angles=True (:150-229), signed_torsion_angle (:231-273)
and periodic_torsion (:276-332). prefactor = lamb_offset + lamb_mult * lamb
(defaults mult 0, offset 1 as in harmonic_bond.cu:41-53).
"""

from __future__ import annotations

import numpy as np
import torch


def _t(a, g=False):
    t = torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=torch.float64).clone()
    t.requires_grad_(g)
    return t


def _prefactor(n, lamb, mult, offset):
    mult = torch.zeros(n, dtype=torch.float64) if mult is None else torch.as_tensor(np.asarray(mult), dtype=torch.float64)
    offset = torch.ones(n, dtype=torch.float64) if offset is None else torch.as_tensor(np.asarray(offset), dtype=torch.float64)
    return offset + mult * lamb, mult


def harmonic_bond_energy(x, params, lamb, idxs, mult=None, offset=None):
    idxs = torch.as_tensor(np.asarray(idxs, dtype=np.int64).reshape(-1, 2))
    pf, _ = _prefactor(len(idxs), lamb, mult, offset)
    d = x[idxs[:, 0]] - x[idxs[:, 1]]
    r = torch.sqrt((d * d).sum(-1))
    kb, b0 = params[:, 0], params[:, 1]
    return torch.sum(pf * kb / 2 * (r - b0) ** 2)


def harmonic_angle_energy(x, params, lamb, idxs, mult=None, offset=None):
    idxs = torch.as_tensor(np.asarray(idxs, dtype=np.int64).reshape(-1, 3))
    pf, _ = _prefactor(len(idxs), lamb, mult, offset)
    vij = x[idxs[:, 0]] - x[idxs[:, 1]]
    vjk = x[idxs[:, 2]] - x[idxs[:, 1]]
    top = (vij * vjk).sum(-1)
    bot = torch.linalg.norm(vij, dim=-1) * torch.linalg.norm(vjk, dim=-1)
    ka, a0 = params[:, 0], params[:, 1]
    return torch.sum(pf * ka / 2 * (top / bot - torch.cos(a0)) ** 2)


def signed_torsion_angle(ci, cj, ck, cl):
    rij = cj - ci
    rkj = cj - ck
    rkl = cl - ck
    n1 = torch.cross(rij, rkj, dim=-1)
    n2 = torch.cross(rkj, rkl, dim=-1)
    y = (torch.cross(n1, n2, dim=-1) * (rkj / torch.linalg.norm(rkj, dim=-1, keepdim=True))).sum(-1)
    xx = (n1 * n2).sum(-1)
    return torch.atan2(y, xx)


def periodic_torsion_energy(x, params, lamb, idxs, mult=None, offset=None):
    idxs = torch.as_tensor(np.asarray(idxs, dtype=np.int64).reshape(-1, 4))
    pf, _ = _prefactor(len(idxs), lamb, mult, offset)
    ang = signed_torsion_angle(x[idxs[:, 0]], x[idxs[:, 1]], x[idxs[:, 2]], x[idxs[:, 3]])
    k, phase, period = params[:, 0], params[:, 1], params[:, 2]
    return torch.sum(pf * k * (1 + torch.cos(period * ang - phase)))


_ENERGY = {"bond": harmonic_bond_energy, "angle": harmonic_angle_energy, "torsion": periodic_torsion_energy}


def bonded(kind, x, params, lamb, idxs, mult=None, offset=None):
    """(du_dx, du_dp, du_dl, u) in FP64 for kind in {'bond', 'angle', 'torsion'}."""
    xt, pt, lt = _t(x, True), _t(params, True), _t(lamb, True)
    u = _ENERGY[kind](xt, pt, lt, idxs, mult, offset)
    if len(np.asarray(idxs)) == 0:
        return np.zeros_like(np.asarray(x, dtype=np.float64)), np.zeros_like(np.asarray(params, dtype=np.float64)), 0.0, 0.0
    gx, gp, gl = torch.autograd.grad(u, (xt, pt, lt), allow_unused=True)
    gl = 0.0 if gl is None else float(gl)
    return gx.numpy(), gp.numpy(), gl, float(u.detach())


def bonded_dl_dp(kind, x, params, lamb, idxs, mult=None, offset=None):
    """d(du/dlambda)/dparams in FP64 (Torch double backward); shaped like params."""
    params = np.asarray(params, dtype=np.float64)
    if len(np.asarray(idxs)) == 0:
        return np.zeros_like(params)
    xt, pt, lt = _t(x), _t(params, True), _t(lamb, True)
    u = _ENERGY[kind](xt, pt, lt, idxs, mult, offset)
    (gl,) = torch.autograd.grad(u, (lt,), create_graph=True)
    if not gl.requires_grad:
        return np.zeros_like(params)
    (gp,) = torch.autograd.grad(gl, (pt,), allow_unused=True)
    return np.zeros_like(params) if gp is None else gp.numpy()


def bonded_energy_fn(kind, idxs, mult=None, offset=None):
    """U(xt, pt, lt) of one bonded kind as a Torch function (params shaped
    [n_terms, n_params]), for oracle.nonbonded.hvp()."""

    def fn(xt, pt, lt):
        return _ENERGY[kind](xt, pt, lt, idxs, mult, offset)

    return fn
