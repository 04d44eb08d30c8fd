"""Differentiable potential energy U(coords, params, box, lam) over the HIP
potentials, as a torch.autograd.Function.

Restates the contract of the reference's JAX custom_jvp wrapper
(timemachine/fe/functional.py:18-76): one custom-op call per evaluation, the
derivative flags chosen from which inputs need gradients, and the
Jacobian-vector product <gradient, tangent>. Here the reverse mode is
implemented (vector-Jacobian product = upstream scalar x gradient), which is
what Torch needs. Box derivatives are not supported, as in the reference.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from timemachine_amd.lib.potentials import SummedPotential


def _np(t) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy().astype(np.float64)
    return np.asarray(t, dtype=np.float64)


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
            None if du_dl is None else torch.tensor(float(du_dl), dtype=torch.float64, device=dev),
        )
        return torch.tensor(float(u), dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g):
        gx, gp, gl = ctx.grads
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
