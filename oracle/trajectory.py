"""FP64 backpropagation through Langevin steps (test infrastructure only).

Differentiates a loss on the du/dlambda values stored along a BAOAB trajectory
(baoab_step of oracle.integrator: integrator.cu:37-69 with cb = -dt/m) with
respect to the potential parameters, by Torch autograd through every step of
the trajectory on the FP64 oracle energies. This is the pathwise derivative
that the reference's design notes describe (attic/docs/paper.tex:27: analytic
derivatives through Langevin dynamics from Hessians and mixed second
derivatives of the energy); the reference's code computes none. The noise is
an input (zero for a deterministic check), so no random stream has to match.

du/dlambda is stored as the engine stores it: at the state at the start of
run step j when (j + 1) % interval == 0.
"""

from __future__ import annotations

from typing import Callable, Sequence

import numpy as np
import torch


def trajectory_loss_grad(
    energy: Callable,
    x0: np.ndarray,
    v0: np.ndarray,
    params: Sequence[np.ndarray],
    lam: float,
    n_steps: int,
    interval: int,
    dt: float,
    ca: float,
    cbs: np.ndarray,
    ccs: np.ndarray,
    dl_bar: np.ndarray,
    noise: np.ndarray | None = None,
    detach_every: int = 0,
):
    """Returns (du_dl [F], grads [one array per params entry], x0_bar, v0_bar)
    for the loss L = sum_f dl_bar[f] du_dl_f, where energy(x, [p...], lam) is a
    Torch function and noise [n_steps, N, 3] (or None: zero). detach_every > 0
    cuts the state's history at the start of every block of that many steps
    (truncated backpropagation; x0_bar and v0_bar are then the first block's)."""
    x = torch.tensor(np.asarray(x0, np.float64), requires_grad=True)
    v = torch.tensor(np.asarray(v0, np.float64), requires_grad=True)
    x_leaf, v_leaf = x, v
    ps = [torch.tensor(np.asarray(p, np.float64), requires_grad=True) for p in params]
    cb = torch.tensor(np.asarray(cbs, np.float64))[:, None]
    cc = torch.tensor(np.asarray(ccs, np.float64))[:, None]
    loss = torch.zeros((), dtype=torch.float64)
    dls = []
    for j in range(n_steps):
        if detach_every > 0 and j > 0 and j % detach_every == 0:
            x = x.detach().requires_grad_(True)
            v = v.detach()
        # a lambda leaf per step: du/dlambda is the partial derivative at this
        # step's coordinates (which depend on lambda through earlier forces)
        lt = torch.tensor(float(lam), dtype=torch.float64, requires_grad=True)
        u = energy(x, ps, lt)
        gx, gl = torch.autograd.grad(u, (x, lt), create_graph=True)
        if (j + 1) % interval == 0:
            f = (j + 1) // interval - 1
            dls.append(float(gl.detach()))
            loss = loss + float(dl_bar[f]) * gl
        z = torch.zeros_like(x) if noise is None else torch.tensor(np.asarray(noise[j], np.float64))
        v_mid = v + cb * gx
        v = ca * v_mid + cc * z
        x = x + 0.5 * dt * (v_mid + v)
    grads = torch.autograd.grad(loss, [*ps, x_leaf, v_leaf], allow_unused=True)
    out = [np.zeros_like(np.asarray(p, np.float64)) if g is None else g.numpy() for g, p in
           zip(grads[:len(ps)], params)]
    xb = grads[len(ps)]
    vb = grads[len(ps) + 1]
    return (np.array(dls), out, np.zeros_like(x0) if xb is None else xb.numpy(),
            np.zeros_like(v0) if vb is None else vb.numpy())


def trajectory_du_dl(energy: Callable, x0, v0, params, lam, n_steps, interval, dt, ca, cbs, ccs, noise=None):
    """The stored du/dlambda values of the same trajectory, forward only (for
    finite differences)."""
    x = torch.tensor(np.asarray(x0, np.float64))
    v = torch.tensor(np.asarray(v0, np.float64))
    ps = [torch.tensor(np.asarray(p, np.float64)) for p in params]
    cb = torch.tensor(np.asarray(cbs, np.float64))[:, None]
    cc = torch.tensor(np.asarray(ccs, np.float64))[:, None]
    dls = []
    for j in range(n_steps):
        xr = x.clone().requires_grad_(True)
        lt = torch.tensor(float(lam), dtype=torch.float64, requires_grad=True)
        gx, gl = torch.autograd.grad(energy(xr, ps, lt), (xr, lt))
        if (j + 1) % interval == 0:
            dls.append(float(gl))
        z = torch.zeros_like(x) if noise is None else torch.tensor(np.asarray(noise[j], np.float64))
        v_mid = v + cb * gx
        v = ca * v_mid + cc * z
        x = x + 0.5 * dt * (v_mid + v)
    return np.array(dls)
