"""Fitting harness: kinetic and transport parameters from a line scan.

The spark's RyR trajectory is frozen (a recorded schedule, LCC states held),
so the simulated dF/F0 line scan is a smooth function of the parameters
theta (srs_amd.adjoint_theta(): D_Ca, D_Fluo, kon/koff of CaM, TnC, ATP and
Fluo-4, RyR unitary flux, SERCA pump density, SL buffer K_D).  The loss
sum (dF/F0 - target)^2 and its gradient come from the library's adjoint
(srs_linescan_loss_grad: checkpoints plus recomputation on the device; for
many realizations, srs_linescan_loss_grad_batch and the batch-mean loss);
LinescanLoss exposes them to torch.autograd.  torch is used here only as the
optimiser's front end; nothing in the simulation path depends on it.
"""
from __future__ import annotations

import numpy as np

from . import Context, Params, adjoint_theta


class LinescanProblem:
    """A one-realization context with a scan line, a frozen schedule, sample
    steps and a target; loss(theta) and grad(theta) at fixed initial state."""

    def __init__(self, ctx: Context, scan_id: int, schedule: np.ndarray, sample_steps, target: np.ndarray,
                 params: Params, X0: np.ndarray, ckpt: int = 256):
        self.ctx, self.scan_id = ctx, scan_id
        self.schedule = np.ascontiguousarray(schedule, np.int32)
        self.sample_steps = np.ascontiguousarray(sample_steps, np.int64)
        self.target = np.ascontiguousarray(target, np.float64)
        self.params = params.copy()
        self.X0 = np.ascontiguousarray(X0, np.float64)
        self.ckpt = ckpt
        self.idx = adjoint_theta()

    def theta(self) -> np.ndarray:
        return self.params.reals[self.idx].copy()

    def _set(self, theta):
        P = self.params.reals.copy()
        P[self.idx] = np.asarray(theta, np.float64)
        self.ctx.params_set(0, P)
        self.ctx.state_set(0, self.X0)

    def loss(self, theta) -> float:
        self._set(theta)
        return self.ctx.loss_grad(0, self.scan_id, self.schedule, self.sample_steps, self.target, self.ckpt,
                                  grad=False)

    def loss_grad(self, theta):
        self._set(theta)
        return self.ctx.loss_grad(0, self.scan_id, self.schedule, self.sample_steps, self.target, self.ckpt)


class BatchLinescanProblem:
    """The batch-mean loss over realizations `reals` of one context, each with
    its own frozen schedule, target and parameter vector; theta is shared
    (set in every member's parameters), the gradient is the mean of the
    members' adjoint gradients (srs_linescan_loss_grad_batch)."""

    def __init__(self, ctx: Context, scan_id: int, reals, schedules, sample_steps, targets: np.ndarray,
                 params: list, X0: list, ckpt: int = 0, max_batch: int = 0):
        self.ctx, self.scan_id = ctx, scan_id
        self.reals = [int(r) for r in reals]
        self.schedules = [np.ascontiguousarray(s, np.int32) for s in schedules]
        self.sample_steps = np.ascontiguousarray(sample_steps, np.int64)
        self.targets = np.ascontiguousarray(targets, np.float64)
        self.params = [p.copy() for p in params]
        self.X0 = [np.ascontiguousarray(x, np.float64) for x in X0]
        self.ckpt, self.max_batch = ckpt, max_batch
        self.idx = adjoint_theta()

    def theta(self) -> np.ndarray:
        return self.params[0].reals[self.idx].copy()

    def _set(self, theta):
        for r, p, x in zip(self.reals, self.params, self.X0):
            P = p.reals.copy()
            P[self.idx] = np.asarray(theta, np.float64)
            self.ctx.params_set(r, P)
            self.ctx.state_set(r, x)

    def losses(self, theta) -> np.ndarray:
        self._set(theta)
        return self.ctx.loss_grad_batch(self.reals, self.scan_id, self.schedules, self.sample_steps, self.targets,
                                        self.ckpt, self.max_batch, grad=False)

    def loss(self, theta) -> float:
        return float(np.mean(self.losses(theta)))

    def loss_grad(self, theta):
        self._set(theta)
        L, g = self.ctx.loss_grad_batch(self.reals, self.scan_id, self.schedules, self.sample_steps, self.targets,
                                        self.ckpt, self.max_batch)
        return float(np.mean(L)), g.mean(axis=0)


def linescan_loss_fn():
    """torch.autograd.Function over a LinescanProblem or BatchLinescanProblem
    (imports torch lazily)."""
    import torch

    class LinescanLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, problem):
            loss, g = problem.loss_grad(theta.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(g))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None

    return LinescanLoss


def fit(problem, steps: int = 20, lr: float = 0.05, log=None) -> np.ndarray:
    """Adam on log(theta) through the autograd wrapper; returns theta."""
    import torch

    Loss = linescan_loss_fn()
    z = torch.tensor(np.log(problem.theta()), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([z], lr=lr)
    for k in range(steps):
        opt.zero_grad()
        loss = Loss.apply(torch.exp(z), problem)
        loss.backward()
        opt.step()
        if log:
            log(f"fit step {k}: loss {loss.item():.6e}")
    return np.exp(z.detach().numpy())
