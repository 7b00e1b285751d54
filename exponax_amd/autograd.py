"""Differentiable stepping for PyTorch: ``torch.autograd.Function`` wrappers
whose backward passes run the hand-written adjoint kernels.

The reference gets its derivatives from ``jax.grad`` through ``lax.scan``
(``_utils.py:92-254``, README.md:85-91). It keeps every step's residuals. Here
the backward of a rollout keeps one state every ``checkpoint_every`` steps
and recomputes each segment on the way back.

    u0 = torch.randn(32, 1, 128, 128, device="cuda", requires_grad=True)
    uT = exponax_amd.autograd.repeat(stepper, u0, 50, checkpoint_every=10)
    loss = ((uT - target) ** 2).mean()
    loss.backward()          # u0.grad via exmi_rollout_vjp
"""
from __future__ import annotations

import torch


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, stepper):
        ctx.stepper = stepper
        ctx.save_for_backward(u)
        return stepper.step(u.detach())

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        return ctx.stepper.step_vjp(u.detach(), g.contiguous()), None


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, stepper, n, include_init, checkpoint_every):
        ctx.stepper, ctx.n, ctx.include_init, ctx.k = stepper, n, include_init, checkpoint_every
        ctx.save_for_backward(u0)
        return stepper.rollout_batch(u0.detach(), n, include_init=include_init)

    @staticmethod
    def backward(ctx, g_traj):
        (u0,) = ctx.saved_tensors
        g = ctx.stepper.rollout_vjp(u0.detach(), ctx.n, g_traj=g_traj.contiguous(), include_init=ctx.include_init,
                                    checkpoint_every=ctx.k)
        return g, None, None, None, None


class _Repeat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, stepper, n, checkpoint_every):
        ctx.stepper, ctx.n, ctx.k = stepper, n, checkpoint_every
        ctx.save_for_backward(u0)
        return stepper.repeat_batch(u0.detach(), n)

    @staticmethod
    def backward(ctx, g):
        (u0,) = ctx.saved_tensors
        return ctx.stepper.rollout_vjp(u0.detach(), ctx.n, g_final=g.contiguous(), checkpoint_every=ctx.k), None, \
            None, None


def _check(u):
    if not (isinstance(u, torch.Tensor) and u.is_cuda):
        raise TypeError("exponax_amd.autograd expects a torch tensor on the HIP device")


def step(stepper, u):
    """One step of a batch ``(..., C, N^D)``, differentiable in ``u``."""
    _check(u)
    return _Step.apply(u, stepper)


def rollout(stepper, u0, n: int, *, include_init: bool = False, checkpoint_every: int = 10):
    """``exponax.rollout(stepper, n, include_init)(u0)`` batched over leading
    axes: ``(..., n[+1], C, N^D)``, differentiable in ``u0``."""
    _check(u0)
    return _Rollout.apply(u0, stepper, n, include_init, checkpoint_every)


def repeat(stepper, u0, n: int, *, checkpoint_every: int = 10):
    """``exponax.repeat(stepper, n)(u0)``, differentiable in ``u0``."""
    _check(u0)
    return _Repeat.apply(u0, stepper, n, checkpoint_every)
