"""Calibration harness for the pathwise adjoint (Python side only).

Observable: the time average over K steps of the mean x-velocity of the
particles in a band of cells [lo, hi) along `axis` (for a channel, the
centreline band), evaluated after each step.  Parameters: the body force
GRAV and the SRD rotation angle RA.

The forward pass runs K steps with the library, saving a device checkpoint
every `ckpt` steps and the band membership of every step.  The backward pass
restores each checkpoint from the last one down, re-runs its steps while the
library records the tape, and pops them with mpcd_gpu_step_backward, feeding
dJ_t/dv at the end of every step.  Cell assignments, draws and wall events are
those of the recorded run (pathwise gradient).

`CentrelineVelocity` wraps the two passes as a torch.autograd.Function of the
parameter vector (g_x, RA).

`ReverseRun` is the same checkpoint-and-recompute scheme for every operator
of the step, the nematic LC collision with Jeffery rotation and dipole
activity included, through mpcd_gpu_step_backward_state (the adjoint of
positions, velocities and orientations): observables are the band velocity
or the mean scalar order parameter of the cells (mpcd_gpu_observable), and
the gradient is taken with respect to the body force and each species'
activity.  `ActiveChannelReverse` wraps it as a torch.autograd.Function of
(g_x, activity).

The forward-mode tangent (mpcd_gpu_tangent) gives the same pathwise
derivative for every operator of the step, including the nematic LC
collision with activity: `tangent_value_and_grad` runs the K steps once per
direction in parameter space (body force, species activity) and averages the
band's tangent velocity beside its velocity.  `ActiveChannel` wraps it as a
torch.autograd.Function of (g_x, activity).
"""
from __future__ import annotations

import numpy as np


class AdjointRun:
    def __init__(self, M, cfg, Q0, V0, K, seed, band, axis=1, ckpt=20, precision=None):
        self.M, self.cfg, self.K, self.seed, self.band, self.axis, self.ckpt = M, cfg, K, seed, band, axis, ckpt
        self.N = int(Q0.shape[0])
        self.ctx = M.Context(cfg, self.N, seed=seed, precision=M.FP64 if precision is None else precision)
        self.ctx.upload(0, Q0, V0)
        self.masks = []
        self.J = None

    def _mask(self, Q):
        y = np.trunc(Q[:, self.axis])
        return (y >= self.band[0]) & (y < self.band[1])

    def forward(self) -> float:
        ctx = self.ctx
        ctx.record(1)  # recording keeps every step on the path the tape uses
        J = 0.0
        for t in range(self.K):
            if t % self.ckpt == 0:
                ctx.checkpoint(t // self.ckpt)
            ctx.step()
            Q, V, _ = ctx.download(0)
            m = self._mask(Q)
            self.masks.append(m)
            J += V[m, 0].mean() / self.K
        self.J = J
        return J

    def backward(self):
        """dJ/dGRAV (3 components) and dJ/dRA"""
        ctx, K = self.ctx, self.K
        ctx.adjoint_reset()
        nseg = (K + self.ckpt - 1) // self.ckpt
        th = np.zeros((1, 4))
        for s in reversed(range(nseg)):
            t0, t1 = s * self.ckpt, min(K, (s + 1) * self.ckpt)
            ctx.restore(s)
            ctx.record(t1 - t0)
            for _ in range(t0, t1):
                ctx.step()
            for t in reversed(range(t0, t1)):
                adj = np.zeros((self.N, 3))
                m = self.masks[t]
                adj[m, 0] = 1.0 / (K * m.sum())
                _, th = ctx.step_backward(adj)
        ctx.record(0)
        return th[0, :3].copy(), float(th[0, 3])

    def close(self):
        self.ctx.close()


class ReverseRun:
    """J = (1/K) sum_t obs(state after step t) and its pathwise gradient in
    (GRAV, ACT) by the reverse mode of the whole step, checkpointing every
    `ckpt` steps.  obs = (kind, arg) of mpcd_gpu_observable."""

    def __init__(self, M, cfg, Q0, V0, K, seed, obs, U0=None, spid=None, ckpt=20):
        self.M, self.K, self.ckpt, self.obs = M, K, ckpt, obs
        self.N = int(Q0.shape[0])
        self.ctx = M.Context(cfg, self.N, seed=seed, precision=M.FP64)
        self.ctx.upload(0, Q0, V0, U=U0, spid=spid)
        self.J = None

    def forward(self) -> float:
        ctx = self.ctx
        ctx.record(1)  # recording keeps every step on the path the replay uses
        J = 0.0
        for t in range(self.K):
            if t % self.ckpt == 0:
                ctx.checkpoint(t // self.ckpt)
            ctx.step()
            v, _ = ctx.observable(*self.obs)
            J += v[0] / self.K
        self.J = J
        return J

    def backward(self):
        """(dJ/dGRAV [3], dJ/dACT [per species]) of replica 0"""
        ctx, K = self.ctx, self.K
        ctx.adjoint_reset()
        nseg = (K + self.ckpt - 1) // self.ckpt
        th = np.zeros((1, self.M.NTHETA))
        for s in reversed(range(nseg)):
            t0, t1 = s * self.ckpt, min(K, (s + 1) * self.ckpt)
            ctx.restore(s)
            ctx.record(t1 - t0)
            grads = []
            for _ in range(t0, t1):
                ctx.step()
                _, g = ctx.observable(*self.obs, grad=True)
                grads.append(g / K)
            for g in reversed(grads):
                _, th = ctx.step_backward_state(g)
        ctx.record(0)
        return th[0, 0:3].copy(), th[0, 4:].copy()

    def close(self):
        self.ctx.close()


def reverse_value_and_grad(M, cfg, Q0, V0, K, seed, obs, U0=None, spid=None, ckpt=20):
    """J, dJ/dGRAV [3], dJ/dACT [per species] by the reverse mode"""
    run = ReverseRun(M, cfg, Q0, V0, K, seed, obs, U0=U0, spid=spid, ckpt=ckpt)
    try:
        J = run.forward()
        dg, dact = run.backward()
    finally:
        run.close()
    return J, dg, dact


def value_and_grad(M, cfg, Q0, V0, K, seed, band, **kw):
    run = AdjointRun(M, cfg, Q0, V0, K, seed, band, **kw)
    J = run.forward()
    dg, da = run.backward()
    run.close()
    return J, dg, da


def tangent_value_and_grad(M, cfg, Q0, V0, K, seed, band, directions, U0=None, spid=None, axis=1):
    """J and dJ/dtheta_k for each direction k, a dict(dgrav=[3], dact=[nspec]):
    one forward run of K steps per direction, the band and the draws of the
    run itself (pathwise derivative)"""
    N = int(Q0.shape[0])
    J, grads = None, []
    for d in directions:
        ctx = M.Context(cfg, N, seed=seed, precision=M.FP64)
        try:
            ctx.upload(0, Q0, V0, U=U0, spid=spid)
            ctx.tangent(True, dgrav=d.get("dgrav"), dact=d.get("dact"))
            j = dj = 0.0
            for _ in range(K):
                ctx.step()
                Q, V, _ = ctx.download(0)
                _, dV, _ = ctx.tangent_download(0)
                y = np.trunc(Q[:, axis])
                m = (y >= band[0]) & (y < band[1])
                j += V[m, 0].mean() / K
                dj += dV[m, 0].mean() / K
        finally:
            ctx.close()
        if J is not None and j != J:
            raise RuntimeError("tangent runs of one problem disagree on J")
        J = j
        grads.append(dj)
    return J, grads


def _torch_function():
    import torch

    class CentrelineVelocity(torch.autograd.Function):
        """J(theta) with theta = (g_x, RA); problem(g, angle) returns an AdjointRun"""

        @staticmethod
        def forward(ctx, theta, problem):
            run = problem(float(theta[0]), float(theta[1]))
            J = run.forward()
            ctx.run = run
            ctx.like = theta
            return theta.new_tensor(J)

        @staticmethod
        def backward(ctx, gJ):
            dg, da = ctx.run.backward()
            ctx.run.close()
            return gJ * ctx.like.new_tensor([dg[0], da]), None

    return CentrelineVelocity


def centreline_velocity(theta, problem):
    """torch entry point: J = centreline_velocity(theta, problem); J.backward()"""
    return _torch_function().apply(theta, problem)


def _torch_active():
    import torch

    class ActiveChannel(torch.autograd.Function):
        """J(theta) with theta = (g_x, activity of every species); problem(g, act)
        returns (M, cfg, Q0, V0, U0, K, seed, band)"""

        @staticmethod
        def forward(ctx, theta, problem):
            M, cfg, Q0, V0, U0, K, seed, band = problem(float(theta[0]), float(theta[1]))
            nsp = len(cfg["species"])
            J, (dg, da) = tangent_value_and_grad(M, cfg, Q0, V0, K, seed, band,
                                                 [dict(dgrav=[1.0, 0.0, 0.0]), dict(dact=[1.0] * nsp)], U0=U0)
            ctx.grad = theta.new_tensor([dg, da])
            return theta.new_tensor(J)

        @staticmethod
        def backward(ctx, gJ):
            return gJ * ctx.grad, None

    return ActiveChannel


def active_channel_velocity(theta, problem):
    """torch entry point: J = active_channel_velocity(theta, problem); J.backward()"""
    return _torch_active().apply(theta, problem)


def _torch_active_reverse():
    import torch

    class ActiveChannelReverse(torch.autograd.Function):
        """J(theta) with theta = (g_x, activity of every species) by the reverse
        mode; problem(g, act) returns (M, cfg, Q0, V0, U0, K, seed, obs)"""

        @staticmethod
        def forward(ctx, theta, problem):
            M, cfg, Q0, V0, U0, K, seed, obs = problem(float(theta[0]), float(theta[1]))
            run = ReverseRun(M, cfg, Q0, V0, K, seed, obs, U0=U0)
            ctx.run = run
            ctx.nsp = len(cfg["species"])
            ctx.like = theta
            return theta.new_tensor(run.forward())

        @staticmethod
        def backward(ctx, gJ):
            try:
                dg, dact = ctx.run.backward()
            finally:
                ctx.run.close()
            return gJ * ctx.like.new_tensor([dg[0], float(np.sum(dact[:ctx.nsp]))]), None

    return ActiveChannelReverse


def active_channel_reverse(theta, problem):
    """torch entry point (reverse mode): J = active_channel_reverse(theta, problem); J.backward()"""
    return _torch_active_reverse().apply(theta, problem)
