"""Second-order autograd of the differentiable potential interface, on the CPU.

The Torch plumbing of timemachine_amd.fe.functional (double backward through
_Gradient, whose vector-Jacobian product is execute_hvp) is checked with a
stand-in impl whose energy has a known gradient and Hessian:
    U(x, p, lam) = sum_i p_i0 |x_i|^2 + lam^2 sum_i p_i1 + lam sum_i x_i0 p_i1
The same energy written in Torch and differentiated by Torch is the
reference. The HIP potentials behind the same interface are checked against
the FP64 oracle on the GPU (test_hvp_gpu.py).
"""

import numpy as np
import torch

from timemachine_amd.fe import functional


class QuadraticImpl:
    def execute_selective(self, x, p, box, lam, cx, cp, cl, cu):
        gx = 2 * p[:, :1] * x
        gx[:, 0] += lam * p[:, 1]
        gp = np.stack([(x * x).sum(1), lam * lam + lam * x[:, 0]], 1)
        gl = 2 * lam * p[:, 1].sum() + (x[:, 0] * p[:, 1]).sum()
        u = (p[:, 0] * (x * x).sum(1)).sum() + lam * lam * p[:, 1].sum() + lam * (x[:, 0] * p[:, 1]).sum()
        return (gx if cx else None, gp if cp else None, gl if cl else None, u if cu else None)

    def execute_hvp(self, x, p, box, lam, vx, vp, vl, cx=True, cp=True, cl=True):
        vx = np.zeros_like(x) if vx is None else vx
        vp = np.zeros_like(p) if vp is None else vp
        hx = 2 * p[:, :1] * vx + 2 * vp[:, :1] * x
        hx[:, 0] += vl * p[:, 1] + lam * vp[:, 1]
        hp = np.stack([(2 * x * vx).sum(1), (2 * lam + x[:, 0]) * vl + lam * vx[:, 0]], 1)
        hl = 2 * vl * p[:, 1].sum() + 2 * lam * vp[:, 1].sum() + (vx[:, 0] * p[:, 1] + x[:, 0] * vp[:, 1]).sum()
        return (hx if cx else None, hp if cp else None, float(hl) if cl else None)


def torch_energy(x, p, lam):
    return (p[:, 0] * (x * x).sum(1)).sum() + lam * lam * p[:, 1].sum() + lam * (x[:, 0] * p[:, 1]).sum()


def _inputs(seed=0):
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.normal(size=(5, 3)), dtype=torch.float64, requires_grad=True)
    p = torch.tensor(rng.uniform(0.5, 2.0, size=(5, 2)), dtype=torch.float64, requires_grad=True)
    lam = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    return x, p, lam


def test_first_order_unchanged():
    U = functional.wrap_impl(QuadraticImpl())
    x, p, lam = _inputs()
    u = U(x, p, np.eye(3), lam)
    gx, gp, gl = torch.autograd.grad(u, (x, p, lam))
    rx, rp, rl = torch.autograd.grad(torch_energy(x, p, lam), (x, p, lam))
    for a, b in zip((gx, gp, gl), (rx, rp, rl)):
        torch.testing.assert_close(a, b, rtol=1e-12, atol=1e-12)


def test_force_matching_loss_second_order():
    """loss = |du/dx - F|^2 + (du/dl)^2: its gradient in (x, p, lam) needs
    H_xx, H_xp, H_xl and H_lp, H_ll through the double backward."""
    U = functional.wrap_impl(QuadraticImpl())
    target = torch.tensor(np.random.default_rng(3).normal(size=(5, 3)), dtype=torch.float64)

    def loss(energy, x, p, lam):
        gx, gl = torch.autograd.grad(energy, (x, lam), create_graph=True)
        return ((gx - target) ** 2).sum() + gl ** 2

    x, p, lam = _inputs()
    got = torch.autograd.grad(loss(U(x, p, np.eye(3), lam), x, p, lam), (x, p, lam))
    x2, p2, lam2 = _inputs()
    want = torch.autograd.grad(loss(torch_energy(x2, p2, lam2), x2, p2, lam2), (x2, p2, lam2))
    for a, b in zip(got, want):
        torch.testing.assert_close(a, b, rtol=1e-12, atol=1e-12)


def test_gradient_of_du_dp_loss():
    """A loss on du/dp differentiates through H_px and H_pp."""
    U = functional.wrap_impl(QuadraticImpl())
    x, p, lam = _inputs(5)
    (gp,) = torch.autograd.grad(U(x, p, np.eye(3), lam), (p,), create_graph=True)
    got = torch.autograd.grad((gp ** 2).sum(), (x, p))
    x2, p2, lam2 = _inputs(5)
    (gp2,) = torch.autograd.grad(torch_energy(x2, p2, lam2), (p2,), create_graph=True)
    # U is linear in p here: Torch reports the p part as unused (zero)
    want = torch.autograd.grad((gp2 ** 2).sum(), (x2, p2), allow_unused=True)
    for a, b, ref in zip(got, want, (x2, p2)):
        torch.testing.assert_close(a, torch.zeros_like(ref) if b is None else b, rtol=1e-12, atol=1e-12)


def test_third_order_is_refused():
    """The Hessian-vector product is not itself differentiable: asking for a
    third derivative raises instead of returning zeros."""
    U = functional.wrap_impl(QuadraticImpl())
    x, p, lam = _inputs()
    (gx,) = torch.autograd.grad(U(x, p, np.eye(3), lam), (x,), create_graph=True)
    (hx,) = torch.autograd.grad((gx ** 2).sum(), (x,), create_graph=True)
    try:
        torch.autograd.grad(hx.sum(), (x,))
    except RuntimeError:
        return
    raise AssertionError("third derivative did not raise")
