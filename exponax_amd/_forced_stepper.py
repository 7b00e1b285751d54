"""``ForcedStepper`` (``_forced_stepper.py:8-56``): a forward-Euler forcing
split around any stepper, u -> stepper(u + dt f); the forcing is added on the
GPU (``exmi_axpy``)."""
from __future__ import annotations

from . import _lib
from ._tensor import as_device_tensor, restore_like


def _axpy(x, a, y):
    import torch

    tx, back = as_device_tensor(x)
    ty, _ = as_device_tensor(y)
    if tuple(tx.shape) != tuple(ty.shape):
        raise ValueError(f"state and forcing shapes differ: {tuple(tx.shape)} vs {tuple(ty.shape)}")
    dt = torch.complex64 if (tx.is_complex() or ty.is_complex()) else torch.float32
    tx = tx.to(dt).contiguous()
    ty = ty.to(dt).contiguous()
    out = torch.empty_like(tx)
    n = tx.numel() * (2 if dt == torch.complex64 else 1)
    _lib.check(_lib.lib().exmi_axpy(n, float(a), tx.data_ptr(), ty.data_ptr(), out.data_ptr(), _lib.stream_ptr()),
               "exmi_axpy")
    return restore_like(out, back)


class ForcedStepper:
    def __init__(self, stepper):
        self.stepper = stepper

    def step(self, u, f):
        """``_forced_stepper.py:33-43``."""
        return self.stepper.step(_axpy(u, self.stepper.dt, f))

    def step_fourier(self, u_hat, f_hat):
        """``_forced_stepper.py:45-56``."""
        return self.stepper.step_fourier(_axpy(u_hat, self.stepper.dt, f_hat))

    def __call__(self, u, f):
        return self.step(u, f)


__all__ = ["ForcedStepper"]
