"""FP64 nonbonded potential: LJ 12-6 + erfc-damped Coulomb with 4D decoupling.

Restates timemachine/potentials/nonbonded.py:116-264 (convert_exclusions_to_rescale_masks,
nonbonded_v3) and timemachine/potentials/jax_utils.py:32-148 (convert_to_4d, delta_r,
distance) in Torch FP64; derivatives come from Torch autograd the way the
reference's tests take them from jax.grad (tests/common.py:411-477).

Two evaluators:
  * nonbonded_v3         dense O(N^2), a line-by-line restatement (small N);
  * nonbonded_pairs      the same energy summed over the pair list within the
                         cutoff (large N); pinned to the dense form in tests.
Both take ``interpolated=True`` for parameters [2N, 3] = (initial, final) used
at lambda as (1 - lambda) initial + lambda final (nonbonded.py:341-353).
``interaction_group`` restricts the pair sum to (row atom, other atom) pairs.
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

ONE_4PI_EPS0 = 138.935456


def convert_exclusions_to_rescale_masks(exclusion_idxs, scales, N):
    """nonbonded.py:116-130: mask = 1 - scale on excluded pairs, 1 elsewhere."""
    q_mask = np.ones((N, N))
    lj_mask = np.ones((N, N))
    for (i, j), (sq, slj) in zip(np.asarray(exclusion_idxs), np.asarray(scales)):
        q_mask[i, j] = q_mask[j, i] = 1 - sq
        lj_mask[i, j] = lj_mask[j, i] = 1 - slj
    return q_mask, lj_mask


def _t(a, requires_grad=False):
    t = torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=torch.float64).clone()
    t.requires_grad_(requires_grad)
    return t


def nonbonded_v3_energy(conf, params, box, lamb, q_mask, lj_mask, beta, cutoff, plane, offset):
    """Energy of nonbonded.py:133-264 on torch tensors (conf [N,3], params [N,3], lamb scalar)."""
    N = conf.shape[0]
    plane = torch.as_tensor(np.asarray(plane), dtype=torch.float64)
    offset = torch.as_tensor(np.asarray(offset), dtype=torch.float64)
    w = cutoff * (plane + offset * lamb)  # jax_utils.py:40-55
    x4 = torch.cat([conf, w[:, None]], dim=1)
    box_diag = torch.cat([torch.diagonal(torch.as_tensor(np.asarray(box), dtype=torch.float64)), torch.tensor([1000.0], dtype=torch.float64)])
    diff = x4[None, :, :] - x4[:, None, :]  # delta_r, jax_utils.py:72-77
    diff = diff - box_diag * torch.floor(diff / box_diag + 0.5)
    d2 = (diff * diff).sum(-1)
    eye = torch.eye(N, dtype=torch.bool)
    d2 = torch.where(eye, torch.zeros_like(d2), d2)
    dij = torch.where(eye, torch.zeros_like(d2), torch.sqrt(torch.where(eye, torch.ones_like(d2), d2)))

    q = params[:, 0]
    sig = params[:, 1]
    eps = params[:, 2]
    sig_ij = sig[None, :] + sig[:, None]
    eps_ij = eps[None, :] * eps[:, None]
    keep = (~eye) & (eps_ij != 0)
    eps_ij = torch.where(dij < cutoff, eps_ij, torch.zeros_like(eps_ij))
    sig_ij = torch.where(keep, sig_ij, torch.zeros_like(sig_ij))
    eps_ij = torch.where(keep, eps_ij, torch.zeros_like(eps_ij))
    inv_dij = torch.where(eye, torch.zeros_like(dij), 1.0 / torch.where(eye, torch.ones_like(dij), dij))
    sig2 = sig_ij * inv_dij
    sig2 = sig2 * sig2
    sig6 = sig2 * sig2 * sig2
    e_lj = 4 * eps_ij * (sig6 - 1.0) * sig6
    e_lj = torch.where(keep, e_lj, torch.zeros_like(e_lj))

    qij = q[None, :] * q[:, None]
    qij = torch.where(eye, torch.zeros_like(qij), qij)
    e_q = torch.where(eye, torch.zeros_like(qij), qij * torch.special.erfc(beta * dij) * inv_dij)
    e_q = torch.where(dij > cutoff, torch.zeros_like(e_q), e_q)
    qm = torch.as_tensor(np.asarray(q_mask), dtype=torch.float64)
    ljm = torch.as_tensor(np.asarray(lj_mask), dtype=torch.float64)
    return torch.sum(e_lj * ljm + e_q * qm) / 2


def interpolated(energy_fn):
    """nonbonded.py:341-353: energy over [2N, 3] parameters, the first N rows
    weighted (1 - lambda) and the last N rows lambda."""

    def fn(xt, pt, lt):
        n = pt.shape[0] // 2
        return energy_fn(xt, (1 - lt) * pt[:n] + lt * pt[n:], lt)

    return fn


_interpolated = interpolated  # the evaluators' keyword argument of the same name shadows it


def _grads(energy_fn, x, params, lamb):
    xt = _t(x, True)
    pt = _t(params, True)
    lt = _t(lamb, True)
    u = energy_fn(xt, pt, lt)
    gx, gp, gl = torch.autograd.grad(u, (xt, pt, lt), allow_unused=True)
    gp = torch.zeros_like(pt) if gp is None else gp
    gl = torch.zeros_like(lt) if gl is None else gl
    return gx.numpy(), gp.numpy(), float(gl), float(u.detach())


def nonbonded_v3(x, params, box, lamb, exclusion_idxs, scales, beta, cutoff, plane, offset, interpolated=False):
    """(du_dx [N,3], du_dp [N,3] ([2N,3] interpolated), du_dl, u) of the dense FP64 restatement."""
    N = x.shape[0]
    qm, ljm = convert_exclusions_to_rescale_masks(exclusion_idxs, scales, N)
    fn = lambda xt, pt, lt: nonbonded_v3_energy(xt, pt, box, lt, qm, ljm, beta, cutoff, plane, offset)  # noqa: E731
    return _grads(_interpolated(fn) if interpolated else fn, x, params, lamb)


def candidate_pairs(x: np.ndarray, box: np.ndarray, r: float) -> np.ndarray:
    """All pairs i < j with 3D periodic distance < r (superset of the 4D set)."""
    from scipy.spatial import cKDTree

    L = np.diag(box).astype(np.float64)
    xw = np.mod(x, L)
    xw = np.where(xw >= L, 0.0, xw)
    tree = cKDTree(xw, boxsize=L)
    pairs = tree.query_pairs(r, output_type="ndarray")
    if len(pairs) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    pairs = np.sort(pairs, axis=1)
    return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]


def _pair_energy(xt, pt, lt, pairs, box_diag, beta, cutoff, plane, offset, scale_q=None, scale_lj=None):
    i = torch.as_tensor(pairs[:, 0])
    j = torch.as_tensor(pairs[:, 1])
    w = cutoff * (plane + offset * lt)
    d = xt[i] - xt[j]
    d = d - box_diag * torch.floor(d / box_diag + 0.5)
    dw = w[i] - w[j]
    d2 = (d * d).sum(-1) + dw * dw
    r = torch.sqrt(d2)
    inside = r < cutoff
    q = pt[:, 0]
    sig = pt[:, 1]
    eps = pt[:, 2]
    eps_ij = eps[i] * eps[j]
    lj_on = inside & (eps_ij != 0)
    sig_ij = torch.where(lj_on, sig[i] + sig[j], torch.zeros_like(r))
    eps_ij = torch.where(lj_on, eps_ij, torch.zeros_like(r))
    rs = torch.where(inside, r, torch.ones_like(r))
    s6 = (sig_ij / rs) ** 6
    e_lj = torch.where(lj_on, 4 * eps_ij * (s6 - 1.0) * s6, torch.zeros_like(r))
    e_q = torch.where(inside, q[i] * q[j] * torch.special.erfc(beta * rs) / rs, torch.zeros_like(r))
    if scale_q is not None:
        e_q = e_q * scale_q
        e_lj = e_lj * scale_lj
    return e_lj.sum() + e_q.sum()


def nonbonded_pairs(
    x, params, box, lamb, exclusion_idxs, scales, beta, cutoff, plane, offset, pairs: Optional[np.ndarray] = None,
    interpolated: bool = False,
) -> Tuple[np.ndarray, np.ndarray, float, float]:
    """Same energy as nonbonded_v3 summed over pairs (all pairs within the cutoff,
    minus scale * pair term for each excluded pair)."""
    x = np.asarray(x, dtype=np.float64)
    if pairs is None:
        pairs = candidate_pairs(x, box, cutoff)
    box_diag = torch.tensor(np.diag(np.asarray(box)).copy(), dtype=torch.float64)
    plane_t = torch.as_tensor(np.asarray(plane), dtype=torch.float64)
    offset_t = torch.as_tensor(np.asarray(offset), dtype=torch.float64)
    excl = np.asarray(exclusion_idxs, dtype=np.int64).reshape(-1, 2)
    sc = torch.as_tensor(np.asarray(scales, dtype=np.float64).reshape(-1, 2))

    def fn(xt, pt, lt):
        u = _pair_energy(xt, pt, lt, pairs, box_diag, beta, cutoff, plane_t, offset_t)
        if len(excl):
            u = u - _pair_energy(xt, pt, lt, excl, box_diag, beta, cutoff, plane_t, offset_t, sc[:, 0], sc[:, 1])
        return u

    return _grads(_interpolated(fn) if interpolated else fn, x, params, lamb)


def interaction_group(
    x, params, box, lamb, row_idxs, beta, cutoff, plane, offset, interpolated: bool = False
) -> Tuple[np.ndarray, np.ndarray, float, float]:
    """nonbonded_v3_interaction_groups (nonbonded.py:318-338, on the pairs of
    nonbonded_v3_on_specific_pairs, :266-315) over the 4D coordinates, as
    tests/nonbonded/test_nonbonded_interaction_group.py:93-104 builds them:
    every pair (i, j) with i a row atom and j any other atom, within the cutoff.
    Returns (du_dx, du_dp, du_dl, u) like nonbonded_pairs."""
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    is_row = np.zeros(N, dtype=bool)
    is_row[np.asarray(row_idxs, dtype=np.int64).reshape(-1)] = True
    pairs = candidate_pairs(x, box, cutoff)  # 3D within the cutoff: a superset of the 4D set
    pairs = pairs[is_row[pairs[:, 0]] != is_row[pairs[:, 1]]]
    box_diag = torch.tensor(np.diag(np.asarray(box)).copy(), dtype=torch.float64)
    plane_t = torch.as_tensor(np.asarray(plane), dtype=torch.float64)
    offset_t = torch.as_tensor(np.asarray(offset), dtype=torch.float64)

    def fn(xt, pt, lt):
        return _pair_energy(xt, pt, lt, pairs, box_diag, beta, cutoff, plane_t, offset_t)

    return _grads(_interpolated(fn) if interpolated else fn, x, params, lamb)


def _dl_dp(energy_fn, x, params, lamb):
    """d(dU/dlambda)/dparams by Torch FP64 double backward (the mixed second
    derivative jax.grad(jax.grad(U, 3), 1) of the reference's potentials)."""
    xt = _t(x)
    pt = _t(params, True)
    lt = _t(lamb, True)
    u = energy_fn(xt, pt, lt)
    (gl,) = torch.autograd.grad(u, (lt,), create_graph=True)
    if not gl.requires_grad:
        return np.zeros_like(np.asarray(params, dtype=np.float64))
    (gp,) = torch.autograd.grad(gl, (pt,), allow_unused=True)
    return np.zeros_like(pt.detach().numpy()) if gp is None else gp.numpy()


def nonbonded_pairs_dl_dp(
    x, params, box, lamb, exclusion_idxs, scales, beta, cutoff, plane, offset, pairs: Optional[np.ndarray] = None
) -> np.ndarray:
    """d(du/dlambda)/dp [N, 3] of the energy of nonbonded_pairs (exclusions included)."""
    x = np.asarray(x, dtype=np.float64)
    if pairs is None:
        pairs = candidate_pairs(x, box, cutoff)
    box_diag = torch.tensor(np.diag(np.asarray(box)).copy(), dtype=torch.float64)
    plane_t = torch.as_tensor(np.asarray(plane), dtype=torch.float64)
    offset_t = torch.as_tensor(np.asarray(offset), dtype=torch.float64)
    excl = np.asarray(exclusion_idxs, dtype=np.int64).reshape(-1, 2)
    sc = torch.as_tensor(np.asarray(scales, dtype=np.float64).reshape(-1, 2))

    def fn(xt, pt, lt):
        u = _pair_energy(xt, pt, lt, pairs, box_diag, beta, cutoff, plane_t, offset_t)
        if len(excl):
            u = u - _pair_energy(xt, pt, lt, excl, box_diag, beta, cutoff, plane_t, offset_t, sc[:, 0], sc[:, 1])
        return u

    return _dl_dp(fn, x, params, lamb)


def hvp(energy_fn, x, params, lamb, vx=None, vp=None, vl=0.0):
    """H (vx, vp, vl) with H the Hessian of U in (x, params, lambda), by Torch
    FP64 double backward: the gradient of <grad U, v> (the reference takes such
    second derivatives from jax.grad of its potentials; the trajectory adjoint
    of attic/docs/paper.tex:27 needs them). Returns (hx, hp, hl)."""
    xt, pt, lt = _t(x, True), _t(params, True), _t(lamb, True)
    u = energy_fn(xt, pt, lt)
    gx, gp, gl = torch.autograd.grad(u, (xt, pt, lt), create_graph=True, allow_unused=True)
    s = torch.zeros((), dtype=torch.float64)
    if vx is not None and gx is not None:
        s = s + (gx * _t(vx).reshape(gx.shape)).sum()
    if vp is not None and gp is not None:
        s = s + (gp * _t(vp).reshape(gp.shape)).sum()
    if vl and gl is not None:
        s = s + gl * float(vl)
    if not s.requires_grad:
        return np.zeros_like(xt.detach().numpy()), np.zeros_like(pt.detach().numpy()), 0.0
    hx, hp, hl = torch.autograd.grad(s, (xt, pt, lt), allow_unused=True)
    hx = np.zeros_like(xt.detach().numpy()) if hx is None else hx.numpy()
    hp = np.zeros_like(pt.detach().numpy()) if hp is None else hp.numpy()
    return hx, hp, 0.0 if hl is None else float(hl)


def pair_energy_fn(x, box, exclusion_idxs, scales, beta, cutoff, plane, offset, pairs=None):
    """U(xt, pt, lt) of nonbonded_pairs (all pairs within the cutoff of x, minus
    the scaled excluded pairs) as a Torch function, for hvp()."""
    x = np.asarray(x, dtype=np.float64)
    if pairs is None:
        pairs = candidate_pairs(x, box, cutoff)
    box_diag = torch.tensor(np.diag(np.asarray(box)).copy(), dtype=torch.float64)
    plane_t = torch.as_tensor(np.asarray(plane), dtype=torch.float64)
    offset_t = torch.as_tensor(np.asarray(offset), dtype=torch.float64)
    excl = np.asarray(exclusion_idxs, dtype=np.int64).reshape(-1, 2)
    sc = torch.as_tensor(np.asarray(scales, dtype=np.float64).reshape(-1, 2))

    def fn(xt, pt, lt):
        u = _pair_energy(xt, pt, lt, pairs, box_diag, beta, cutoff, plane_t, offset_t)
        if len(excl):
            u = u - _pair_energy(xt, pt, lt, excl, box_diag, beta, cutoff, plane_t, offset_t, sc[:, 0], sc[:, 1])
        return u

    return fn


def pair_list_energy_fn(pairs, scales, box, beta, cutoff, plane, offset, negated=False):
    """U(xt, pt, lt) of a pair list with per-pair (charge, LJ) scales: the
    scaled pair terms of the listed pairs within the cutoff, negated on request
    (nonbonded.py:266-315 on the listed pairs)."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    box_diag = torch.tensor(np.diag(np.asarray(box)).copy(), dtype=torch.float64)
    plane_t = torch.as_tensor(np.asarray(plane), dtype=torch.float64)
    offset_t = torch.as_tensor(np.asarray(offset), dtype=torch.float64)
    sc = torch.as_tensor(np.asarray(scales, dtype=np.float64).reshape(-1, 2))
    sign = -1.0 if negated else 1.0

    def fn(xt, pt, lt):
        return sign * _pair_energy(xt, pt, lt, pairs, box_diag, beta, cutoff, plane_t, offset_t, sc[:, 0], sc[:, 1])

    return fn


def force_rel_err(test: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """Per-atom |dF| / max(|F_ref|, 1) (tests/common.py:386-409, the OpenMM convention)."""
    num = np.linalg.norm(np.asarray(test) - np.asarray(ref), axis=-1)
    den = np.maximum(np.linalg.norm(np.asarray(ref), axis=-1), 1.0)
    return num / den
