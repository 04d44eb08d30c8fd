"""Differentiable potential energy U(coords, params, box, lam) over the HIP
potentials, as a torch.autograd.Function.

Restates the contract of the reference's JAX custom_jvp wrapper
(timemachine/fe/functional.py:18-76): one custom-op call per evaluation, the
derivative flags chosen from which inputs need gradients, and the
Jacobian-vector product <gradient, tangent>. Here the reverse mode is
implemented (vector-Jacobian product = upstream scalar x gradient), which is
what Torch needs. Box derivatives are not supported, as in the reference.

The energy is twice differentiable: under ``create_graph=True`` the gradient
(du_dx, du_dp, du_dl) is itself a differentiable function of (coords, params,
lam) whose vector-Jacobian product is the potential's Hessian-vector product
(``execute_hvp``, the du_dx adjoint). Losses on forces or on du/dlambda then
backpropagate to the parameters, as jax.grad of the reference's potentials
allows (the reference's custom op itself stops at first order).
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from timemachine_amd.lib.potentials import SummedPotential


def _np(t) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy().astype(np.float64)
    return np.asarray(t, dtype=np.float64)


class _Gradient(torch.autograd.Function):
    """(du_dx, du_dp, du_dl), already evaluated, as a differentiable function
    of (coords, params, lam): the vector-Jacobian product with a cotangent
    (vx, vp, vl) is H (vx, vp, vl), H the symmetric Hessian of U in
    (coords, params, lambda), one execute_hvp call."""

    @staticmethod
    def forward(ctx, coords, params, lam, impl, box, gx, gp, gl):
        ctx.save_for_backward(coords, params, lam)
        ctx.impl = impl
        ctx.box = box
        return gx.clone(), gp.clone(), gl.clone()

    @staticmethod
    @once_differentiable
    def backward(ctx, vx, vp, vl):
        coords, params, lam = ctx.saved_tensors
        need_x, need_p, need_l = ctx.needs_input_grad[:3]
        p_shape = params.shape
        hx, hp, hl = ctx.impl.execute_hvp(
            _np(coords), _np(params).reshape(p_shape), _np(ctx.box), float(lam), _np(vx),
            _np(vp).reshape(p_shape) if params.numel() else None, float(vl), need_x, need_p, need_l,
        )
        dev = coords.device

        def t(a, shape=None):
            a = torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=torch.float64, device=dev)
            return a if shape is None else a.reshape(shape)

        return (
            t(hx) if need_x else None,
            t(hp, p_shape) if need_p else None,
            t(hl, lam.shape) if need_l else None,
            None, None, None, None, None,
        )


class _PotentialEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, params, lam, impl, box):
        need_x, need_p, need_l = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        if isinstance(box, torch.Tensor) and box.requires_grad:
            raise RuntimeError("box derivatives not supported!")
        p_shape = params.shape
        du_dx, du_dp, du_dl, u = impl.execute_selective(
            _np(coords), _np(params).reshape(p_shape), _np(box), float(lam), need_x, need_p, need_l, True
        )
        dev = coords.device if isinstance(coords, torch.Tensor) else torch.device("cpu")
        ctx.grads = (
            None if du_dx is None else torch.as_tensor(du_dx, dtype=torch.float64, device=dev),
            None if du_dp is None else torch.as_tensor(np.asarray(du_dp).reshape(p_shape), dtype=torch.float64, device=dev),
            None if du_dl is None else torch.tensor(float(du_dl), dtype=torch.float64, device=dev).reshape(lam.shape),
        )
        ctx.save_for_backward(coords, params, lam)
        ctx.impl = impl
        ctx.box = box
        return torch.tensor(float(u), dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g):
        gx, gp, gl = ctx.grads
        if torch.is_grad_enabled():
            # create_graph: the gradient is differentiable through _Gradient
            coords, params, lam = ctx.saved_tensors
            zx = torch.zeros_like(coords, dtype=torch.float64) if gx is None else gx
            zp = torch.zeros_like(params, dtype=torch.float64) if gp is None else gp
            zl = torch.zeros_like(lam, dtype=torch.float64) if gl is None else gl
            dx, dp, dl = _Gradient.apply(coords, params, lam, ctx.impl, ctx.box, zx, zp, zl)
            gx = None if gx is None else dx
            gp = None if gp is None else dp
            gl = None if gl is None else dl
        return (
            None if gx is None else g * gx,
            None if gp is None else g * gp,
            None if gl is None else g * gl,
            None,
            None,
        )


def wrap_impl(impl, pack=lambda x: x):
    """U(coords, params, box, lam) -> scalar tensor for one unbound custom-op
    potential (fe/functional.py:18-76). ``pack`` flattens a list of parameter
    arrays into the impl's parameter vector."""

    def U(coords, params, box, lam):
        coords = torch.as_tensor(coords, dtype=torch.float64) if not isinstance(coords, torch.Tensor) else coords
        params = pack(params)
        params = torch.as_tensor(params, dtype=torch.float64) if not isinstance(params, torch.Tensor) else params
        if not isinstance(lam, torch.Tensor):
            lam = torch.tensor(float(lam), dtype=torch.float64)
        return _PotentialEnergy.apply(coords, params, lam, impl, box)

    return U


def construct_differentiable_interface(unbound_potentials, precision=np.float32):
    """Sum of the component potentials evaluated one by one
    (fe/functional.py:79-95)."""
    impls = [ubp.unbound_impl(precision) for ubp in unbound_potentials]
    U_s = [wrap_impl(impl) for impl in impls]

    def U(coords, params: Sequence, box, lam):
        return torch.stack([U_i(coords, p_i, box, lam) for (U_i, p_i) in zip(U_s, params)]).sum()

    return U


def construct_differentiable_interface_fast(unbound_potentials, params, precision=np.float32):
    """Sum of the component potentials in one SummedPotential custom op
    (fe/functional.py:98-119)."""
    impl = SummedPotential(unbound_potentials, params).unbound_impl(precision)

    def pack(ps):
        if isinstance(ps, torch.Tensor):
            return ps.reshape(-1)
        if all(isinstance(p, torch.Tensor) for p in ps):
            return torch.cat([p.reshape(-1) for p in ps])
        return np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1) for p in ps])

    return wrap_impl(impl, pack)
