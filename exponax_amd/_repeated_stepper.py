"""``RepeatedStepper``: ``num_sub_steps`` sub-steps of a stepper per call
(exponax/_repeated_stepper.py:9-143).

The reference transforms once, repeats ``step_fourier`` and transforms back
(``_repeated_stepper.py:56-99``). Here:
- 1-D ETDRK2 steppers sub-step with the Fourier-resident repeat kernel: one
  launch keeps each trajectory's spectrum on chip for all sub-steps, the same
  transform-once sequence (complex linear operators included: the spectrum's
  imaginary Nyquist part is carried from sub-step to sub-step);
- other 1-D plans (orders other than 2) and 2-D and 3-D steppers transform
  once and launch ``step_fourier`` per sub-step;
- 1-D channel-coupled reactions, which have no Fourier-space step, repeat
  the physical step.
"""
from __future__ import annotations

import copy

from ._tensor import as_device_tensor, restore_like


class RepeatedStepper:
    def __init__(self, stepper, num_sub_steps: int):
        if int(num_sub_steps) != num_sub_steps or num_sub_steps < 0:
            raise ValueError(f"num_sub_steps must be a non-negative integer, got {num_sub_steps}")
        self.stepper = stepper
        self.num_sub_steps = int(num_sub_steps)
        self.dt = stepper.dt * num_sub_steps
        self.num_spatial_dims = stepper.num_spatial_dims
        self.domain_extent = stepper.domain_extent
        self.num_points = stepper.num_points
        self.num_channels = stepper.num_channels
        self.dx = stepper.dx
        self._sub = self._spectral_twin(stepper)

    @staticmethod
    def _fourier_resident_ok(st):
        """The Fourier-resident kernels are built for 1-D ETDRK2 plans of the
        fused sizes with an uncoupled nonlinearity (complex linear operators:
        N <= 1024)."""
        import numpy as np

        from . import _lib

        nl = getattr(st, "nonlinear_fun", None)
        coupled = getattr(nl, "kind", None) in (_lib.EXMI_NL_GRAY_SCOTT, _lib.EXMI_NL_BELOUSOV_ZHABOTINSKY)
        cplx = np.iscomplexobj(getattr(st, "linear_operator", np.zeros(1)))
        n = st.num_points
        fused = (n & (n - 1)) == 0 and 16 <= n <= 4096 or n in (48, 50)
        return (st.num_spatial_dims == 1 and st.order == 2 and not coupled and fused
                and n <= (1024 if cplx else 4096))

    @staticmethod
    def _spectral_twin(st):
        """A copy of `st` whose plans keep the spectral state on chip between
        sub-steps (1-D ETDRK2, uncoupled), else `st` itself."""
        if RepeatedStepper._fourier_resident_ok(st):
            if getattr(st, "fourier_resident", False):
                return st
            twin = copy.copy(st)
            twin.fourier_resident = True
            twin._plans = {}
            return twin
        return st

    @property
    def state_shape(self):
        return self.stepper.state_shape

    @property
    def spectral_shape(self):
        return self.stepper.spectral_shape

    # --------------------------------------------------------------- calls
    def step_fourier(self, u_hat):
        """``_repeated_stepper.py:84-99``: repeat ``step_fourier``."""
        for _ in range(self.num_sub_steps):
            u_hat = self.stepper.step_fourier(u_hat)
        return u_hat

    def step(self, u):
        """``_repeated_stepper.py:56-82`` on one state or a batch with leading axes."""
        import torch

        if self.num_sub_steps == 0:
            t, back = as_device_tensor(u)
            return restore_like(t.clone(), back)
        D = self.num_spatial_dims
        nl = getattr(self.stepper, "nonlinear_fun", None)
        from . import _lib

        coupled = getattr(nl, "kind", None) in (_lib.EXMI_NL_GRAY_SCOTT, _lib.EXMI_NL_BELOUSOV_ZHABOTINSKY)
        if (D == 1 and not coupled and not self._fourier_resident_ok(self.stepper)) or D >= 2:
            from .spectral import fft, ifft

            t, back = as_device_tensor(u)
            uh = fft(t.to(torch.float32), num_spatial_dims=D)
            return restore_like(ifft(self.step_fourier(uh), num_spatial_dims=D, num_points=self.num_points), back)
        return self._sub.repeat_batch(u, self.num_sub_steps)

    def __call__(self, u):
        """``_repeated_stepper.py:101-143``: exactly one state."""
        shape = tuple(getattr(u, "shape", ()))
        if shape != self.state_shape:
            raise ValueError(f"Expected shape {self.state_shape}, got {shape}. For batched operation use "
                             "`exponax_amd.vmap` on this function.")
        return self.step(u)

    # ------------------------------------------------- rollouts of the wrapper
    def repeat_batch(self, u0, n: int):
        """n calls of the repeated stepper; each returns to physical space,
        as a reference ``repeat`` of the wrapper does."""
        u = u0
        for _ in range(n):
            u = self.step(u)
        return u

    def rollout_batch(self, u0, n: int, *, include_init: bool = False):
        """(..., C, N^D) -> (..., n[+1], C, N^D): the state after every call."""
        import torch

        t, back = as_device_tensor(u0)
        lead = tuple(t.shape[:t.dim() - len(self.state_shape)])
        if tuple(t.shape[len(lead):]) != tuple(self.state_shape):
            raise ValueError(f"Expected shape (..., {', '.join(map(str, self.state_shape))}), got {tuple(t.shape)}.")
        x = t.contiguous().to(torch.float32)
        T = n + (1 if include_init else 0)
        out = torch.empty(lead + (T,) + tuple(self.state_shape), dtype=torch.float32, device=x.device)
        k = len(lead)
        idx = lambda i: (slice(None),) * k + (i,)  # noqa: E731
        if include_init:
            out[idx(0)] = x
        u = x
        for i in range(n):
            u = self.step(u)
            out[idx(i + (1 if include_init else 0))] = u
        return restore_like(out, back)


__all__ = ["RepeatedStepper"]
