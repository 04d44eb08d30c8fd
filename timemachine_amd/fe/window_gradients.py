"""Lambda-window averages of du/dlambda and their gradients with respect to
forcefield parameters, behind a torch.autograd.Function.

For a window at lambda sampled from exp(-U(x; p) / kT), with A = du/dlambda,

    d<A>/dp = <dA/dp> - (<A dU/dp> - <A><dU/dp>) / kT.

Both terms are estimated on the trajectory's frames. The frames are never
stored: while the batched engine steps the windows, it evaluates A, dU/dp
(the du_dp kernels) and dA/dp (the d(du/dlambda)/dp kernels) at every
observed frame and keeps four running sums on the device
(BatchedContexts.set_grad_observer). With ``recompute=True`` the forward pass
runs without the observer and the backward pass re-runs the windows from the
same starting state (the checkpoint); fixed-point forces and counter-based
noise make the recomputed trajectory bitwise identical, so the frames the
backward pass observes are exactly the ones the forward pass averaged.

The estimator is the derivative of the reweighted average
sum_f A_f(p) w_f(p) / sum_f w_f(p), w_f = exp(-(U_f(p) - U_f(p0)) / kT), at
p = p0, which is what the finite-difference test checks.

The pathwise alternative (method="adjoint") differentiates the sampled
trajectories themselves: reverse mode through every Langevin step, one
Hessian-vector product of the energy per step, frames recomputed from
checkpoints (the route of attic/docs/paper.tex:27). It is exact for the
finite run at hand; its variance grows with the run length (the trajectory's
sensitivity to its own past grows with its Lyapunov exponent), so it suits
short horizons, and the covariance estimator the long-run ensemble average.

The reference differentiates through JAX (timemachine/fe/functional.py:18-119
wraps the custom ops in a custom_jvp); its free-energy loss pipeline
(timemachine/fe/estimator.py) has no counterpart here beyond TI.
"""

from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from timemachine_amd import testsystems as ts
from timemachine_amd.lib import BOLTZ, LangevinIntegrator, MonteCarloBarostat, custom_ops


def window_gradient(stats: dict, kT: float) -> np.ndarray:
    """d<A>/dp of one window from the observer sums (flat parameter order)."""
    n = stats["n"]
    if n == 0:
        raise RuntimeError("no observed frames")
    mean_a = stats["sum_a"] / n
    mean_g = stats["sum_g"] / n
    mean_ag = stats["sum_ag"] / n
    mean_h = stats["sum_h"] / n
    return mean_h - (mean_ag - mean_a * mean_g) / kT


def trapezoid_weights(lambdas: Sequence[float]) -> np.ndarray:
    """Weights c_w with sum_w c_w f(lambda_w) = trapezoid rule over lambda
    (the TI estimate of timemachine/fe/math_utils.py:4)."""
    lam = np.asarray(lambdas, dtype=np.float64)
    c = np.zeros_like(lam)
    d = np.diff(lam)
    c[:-1] += 0.5 * d
    c[1:] += 0.5 * d
    return c


@dataclasses.dataclass
class FlatLayout:
    """Offsets of each parameter block in the engine's flat parameter space of one system."""

    n_atoms: int
    n_bonds: int
    n_angles: int
    n_torsions: int

    @property
    def size(self) -> int:
        return self.n_atoms * 3 + self.n_bonds * 2 + self.n_angles * 2 + self.n_torsions * 3

    def split(self, flat: np.ndarray) -> dict:
        o = 0
        out = {}
        for name, rows, cols in (
            ("nonbonded", self.n_atoms, 3),
            ("bond", self.n_bonds, 2),
            ("angle", self.n_angles, 2),
            ("torsion", self.n_torsions, 3),
        ):
            out[name] = flat[o : o + rows * cols].reshape(rows, cols)
            o += rows * cols
        return out


class LambdaWindows:
    """W lambda windows of one system, stepped together from one fixed starting
    state (the checkpoint) in a BatchedContexts engine.

    Each window runs ``equil_steps`` unobserved steps, then ``n_frames`` x
    ``frame_interval`` steps with du/dlambda (and, when observed, the gradient
    sums) taken every ``frame_interval`` steps.
    """

    def __init__(
        self,
        system: ts.System,
        lambdas: Sequence[float],
        n_frames: int,
        frame_interval: int,
        equil_steps: int = 0,
        temperature: float = 300.0,
        dt: float = 0.0025,
        friction: float = 1.0,
        barostat_interval: int = 0,
        seed: int = 2024,
        minimize_steps: int = 0,
        adjoint_horizon: int = 0,
    ):
        if n_frames <= 0 or frame_interval <= 0:
            raise ValueError("n_frames and frame_interval must be positive")
        self.system = system
        self.lambdas = np.asarray(lambdas, dtype=np.float64)
        self.n_frames = int(n_frames)
        self.frame_interval = int(frame_interval)
        self.equil_steps = int(equil_steps)
        self.temperature = float(temperature)
        self.kT = BOLTZ * self.temperature
        self.dt = float(dt)
        self.friction = float(friction)
        self.barostat_interval = int(barostat_interval)
        self.seed = int(seed)
        # method="adjoint": reverse sweep truncated to blocks of this many steps (0: the whole run)
        self.adjoint_horizon = int(adjoint_horizon)
        self.layout = FlatLayout(system.N, len(system.bond_idxs), len(system.angle_idxs), len(system.torsion_idxs))
        # checkpoint: the starting state of every window
        self.x0 = np.array(system.x, dtype=np.float64)
        self.box0 = np.array(system.box, dtype=np.float64)
        self.v0 = [ts.maxwell_boltzmann(system.masses, self.temperature, self.seed + w) for w in range(len(lambdas))]
        self.last_device_bytes = 0  # device memory of the last batch built (its engine's arena)
        if minimize_steps > 0:
            # the checkpoint is FIRE-minimised once, at the middle lambda
            mid = float(self.lambdas[len(self.lambdas) // 2])
            bps = [p.bound_impl(np.float32) for p in system.bound_potentials()]
            intg = LangevinIntegrator(self.temperature, self.dt, self.friction, system.masses, self.seed).impl()
            one = custom_ops.BatchedContexts.from_windows(self.x0, self.v0[:1], self.box0, intg, [self.seed], bps,
                                                          [mid])
            one.fire_minimize(n_steps=minimize_steps)
            self.x0 = one.get_x_t(0)
            del one

    def _system_with(self, nb_params: Optional[np.ndarray], torsion_params: Optional[np.ndarray]) -> ts.System:
        s = dataclasses.replace(self.system)
        if nb_params is not None:
            s.nb_params = np.asarray(nb_params, dtype=np.float64).reshape(self.system.nb_params.shape)
        if torsion_params is not None:
            s.torsion_params = np.asarray(torsion_params, dtype=np.float64).reshape(self.system.torsion_params.shape)
        return s

    def batch(self, nb_params=None, torsion_params=None) -> Tuple[custom_ops.BatchedContexts, list]:
        """A fresh batch of all windows at the checkpoint with the given
        parameters, built in one engine (window w: integrator seed
        seed + 7919 w, barostat seed + 1), and a list of objects to keep alive
        while it is used (empty: the engine holds its potentials)."""
        sys_ = self._system_with(nb_params, torsion_params)
        bps = [p.bound_impl(np.float32) for p in sys_.bound_potentials()]
        seeds = [self.seed + 7919 * w for w in range(len(self.lambdas))]
        intg = LangevinIntegrator(self.temperature, self.dt, self.friction, sys_.masses, seeds[0]).impl()
        baro = None
        if self.barostat_interval > 0:
            baro = MonteCarloBarostat(sys_.N, 1.0, self.temperature, sys_.groups, self.barostat_interval,
                                      seeds[0] + 1).impl(bps)
        batch = custom_ops.BatchedContexts.from_windows(self.x0, np.stack(self.v0), self.box0, intg, seeds, bps,
                                                        self.lambdas, barostat=baro,
                                                        barostat_seeds=[sd + 1 for sd in seeds])
        self.last_device_bytes = batch.device_bytes()
        return batch, []

    def checkpoint_interval(self) -> int:
        """Steps between the adjoint's checkpoints: about the square root of the
        run (memory and recomputation both O(sqrt(steps))), a multiple of the
        frame interval."""
        total = self.equil_steps + self.n_frames * self.frame_interval
        k = max(1, int(round(np.sqrt(total) / self.frame_interval)))
        return k * self.frame_interval

    def adjoint_run(self, nb_params=None, torsion_params=None):
        """The trajectory adjoint's forward pass: the equilibration and sampled
        steps in one run with checkpoints (du/dlambda stored every
        frame_interval steps from the start). Returns (batch, du_dl series of the
        sampled frames [W, n_frames]); the batch holds the checkpoints for
        adjoint_gradients."""
        if self.barostat_interval > 0:
            raise RuntimeError("the trajectory adjoint needs constant volume (no barostat)")
        if self.equil_steps % self.frame_interval != 0:
            raise RuntimeError("the trajectory adjoint needs equil_steps to be a multiple of frame_interval")
        batch, _ = self.batch(nb_params, torsion_params)
        total = self.equil_steps + self.n_frames * self.frame_interval
        du_dl = batch.adjoint_forward(total, self.frame_interval, self.checkpoint_interval())
        return batch, du_dl[:, self.equil_steps // self.frame_interval:]

    def adjoint_gradients(self, batch, grad_means: np.ndarray) -> np.ndarray:
        """Pathwise gradient (flat parameters of one system) of
        sum_w grad_means[w] <du/dlambda>_w through the trajectories of the last
        adjoint_run: the windows share the parameters, so their gradients add."""
        e = self.equil_steps // self.frame_interval
        W = len(self.lambdas)
        dl_bar = np.zeros((W, e + self.n_frames))
        dl_bar[:, e:] = np.asarray(grad_means, dtype=np.float64)[:, None] / self.n_frames
        grads = batch.adjoint_backward(dl_bar, horizon=self.adjoint_horizon)
        return np.sum(grads, axis=0)

    def sample(self, nb_params=None, torsion_params=None, observe: bool = True):
        """(mean du/dlambda per window [W], du/dlambda series [W, n_frames],
        observer sums per window or None)."""
        batch, keep = self.batch(nb_params, torsion_params)
        if self.equil_steps:
            batch.multiple_steps(self.equil_steps)
        if observe:
            batch.set_grad_observer(self.frame_interval)
        du_dl, _, _ = batch.multiple_steps(self.n_frames * self.frame_interval,
                                           store_du_dl_interval=self.frame_interval)
        stats = [batch.grad_stats(w) for w in range(len(self.lambdas))] if observe else None
        del batch, keep
        return du_dl.mean(axis=1), du_dl, stats


class WindowMeanDuDl(torch.autograd.Function):
    """<du/dlambda>_w for every window as a differentiable function of the
    nonbonded parameters [N, 3] (q, sig/2, sqrt(eps) convention) and the
    torsion force constants [T].

    method "covariance": the thermodynamic gradient above (frames observed on
    the device; with recompute the backward pass re-runs the windows).
    method "adjoint": the pathwise derivative of the trajectory averages,
    reverse mode through every Langevin step (Engine trajectory adjoint: one
    Hessian-vector product per step, frames recomputed from checkpoints). With
    recompute the forward pass is a plain run and the backward pass repeats it
    with checkpoints; without, the forward pass keeps the checkpoints on the
    device until the backward pass. Constant volume only."""

    @staticmethod
    def forward(ctx, nb_params, torsion_k, windows: LambdaWindows, recompute: bool = True,
                method: str = "covariance"):
        if method not in ("covariance", "adjoint"):
            raise ValueError(f"unknown gradient method {method!r}")
        nb = nb_params.detach().cpu().numpy().astype(np.float64)
        tp = np.array(windows.system.torsion_params, dtype=np.float64)
        tp[:, 0] = torsion_k.detach().cpu().numpy()
        ctx.windows = windows
        ctx.params = (nb, tp)
        ctx.method = method
        ctx.stats = None
        ctx.batch = None
        if method == "adjoint" and not recompute:
            ctx.batch, du_dl = windows.adjoint_run(nb, tp)
            means = du_dl.mean(axis=1)
        else:
            means, _, stats = windows.sample(nb, tp, observe=(method == "covariance" and not recompute))
            ctx.stats = stats
        return torch.as_tensor(means, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        windows: LambdaWindows = ctx.windows
        go = grad_out.detach().cpu().numpy()
        if ctx.method == "adjoint":
            batch = ctx.batch
            if batch is None:
                batch, _ = windows.adjoint_run(*ctx.params)  # recompute with checkpoints
            total = windows.adjoint_gradients(batch, go)
            ctx.batch = None
            del batch
        else:
            stats = ctx.stats
            if stats is None:
                # recompute the frames from the checkpoint, observing this time
                _, _, stats = windows.sample(*ctx.params, observe=True)
            total = np.zeros(windows.layout.size)
            for w, st in enumerate(stats):
                total += go[w] * window_gradient(st, windows.kT)
        parts = windows.layout.split(total)
        g_nb = torch.as_tensor(parts["nonbonded"].copy(), dtype=torch.float64)
        g_tk = torch.as_tensor(parts["torsion"][:, 0].copy(), dtype=torch.float64)
        return g_nb, g_tk, None, None, None


def ti_loss(nb_params: torch.Tensor, torsion_k: torch.Tensor, windows: LambdaWindows, target: float,
            recompute: bool = True, method: str = "covariance") -> torch.Tensor:
    """L = (sum_w c_w <du/dlambda>_w - target)^2 with trapezoid weights c_w."""
    means = WindowMeanDuDl.apply(nb_params, torsion_k, windows, recompute, method)
    c = torch.as_tensor(trapezoid_weights(windows.lambdas), dtype=torch.float64)
    return (torch.sum(c * means) - target) ** 2
