"""``BaseStepper`` with exponax's constructor and call interface, running the
fused gfx950 ETDRK kernels.

Follows ``exponax/_base_stepper.py:17-271``: subclasses implement
``_build_linear_operator(derivative_operator)`` (returning the diagonal L of
shape (1 | C, ..., N//2+1)) and ``_build_nonlinear_fun(derivative_operator)``
(returning a descriptor from ``exponax_amd.nonlin_fun``). L is evaluated on
the host in float64 once; the exponential and phi-coefficient tables are
formed in float64 on the device when the first plan is built.

Batching: ``__call__`` accepts exactly one state (C, N, ..., N) like the
reference; ``exponax_amd.vmap(stepper)`` (or ``stepper.step`` on a tensor
with leading batch axes) runs a whole batch in one launch.

Plan options, set as attributes before the first call:
- ``exact_form = True`` evaluates the non-conservative convection with its
  two inverse transforms, as the reference does (EXMI_FLAG_EXACT_FORM).
- ``fourier_resident``: how 1-D ETDRK2 rollouts and repeats carry the state
  from step to step. ``None`` (default): the plan keeps each trajectory's
  spectrum on chip and writes its inverse transform where the reference's
  per-step forward transform of that state would return the same spectrum
  exactly (real linear operator, Nyquist mode removed from the nonlinear term
  by the dealias mask), else it re-derives the spectrum from the written
  physical state every step as the reference does. ``True`` forces the
  spectrum-resident form (EXMI_FLAG_FOURIER_RESIDENT: the Fourier-space
  rollout of the reference's performance notes), ``False`` the per-step
  physical round trip (EXMI_FLAG_PHYSICAL_ROLLOUT). The forms differ by
  float32 rounding only where the default picks the resident one.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import spectral as sp
from ._tensor import as_device_tensor, restore_like


class BaseStepper:
    def __init__(self, num_spatial_dims: int, domain_extent: float, num_points: int, dt: float, *,
                 num_channels: int, order: int, num_circle_points: int = 16, circle_radius: float = 1.0):
        self.num_spatial_dims = num_spatial_dims
        self.domain_extent = domain_extent
        self.num_points = num_points
        self.num_channels = num_channels
        self.dt = dt
        self.dx = domain_extent / num_points
        self.order = order
        self.num_circle_points = num_circle_points
        self.circle_radius = circle_radius

        derivative_operator = sp.build_derivative_operator(num_spatial_dims, domain_extent, num_points)
        lin = np.asarray(self._build_linear_operator(derivative_operator))
        single = (1,) + sp.wavenumber_shape(num_spatial_dims, num_points)
        multi = (num_channels,) + sp.wavenumber_shape(num_spatial_dims, num_points)
        if lin.shape not in (single, multi):
            # _base_stepper.py:111-116
            raise ValueError(f"Expected linear operator to have shape {single} or {multi}, got {lin.shape}.")
        if np.iscomplexobj(lin):
            # odd derivatives (advection, dispersion) make L complex; an
            # imaginary part that is zero to rounding is dropped
            scale = max(1.0, float(np.max(np.abs(lin))))
            if np.max(np.abs(lin.imag)) <= 1e-12 * scale:
                lin = lin.real
        self.linear_operator = np.ascontiguousarray(lin, dtype=np.complex128 if np.iscomplexobj(lin) else np.float64)
        if order not in (0, 1, 2, 3, 4):
            raise NotImplementedError(f"Order {order} not implemented.")  # _base_stepper.py:156-157
        self.nonlinear_fun = self._build_nonlinear_fun(derivative_operator)
        self._plans = {}

    # ------------------------------------------------------------ subclass API
    def _build_linear_operator(self, derivative_operator):
        raise NotImplementedError

    def _build_nonlinear_fun(self, derivative_operator):
        raise NotImplementedError

    # --------------------------------------------------------------- plans
    def _desc(self, batch: int, lin_batch: int = 1) -> _lib.PlanDesc:
        d = _lib.PlanDesc()
        d.num_spatial_dims = self.num_spatial_dims
        d.num_points = self.num_points
        d.num_channels = self.num_channels
        d.batch = batch
        d.order = self.order
        d.lin_batch = lin_batch
        d.lin_channels = self.linear_operator.shape[0]
        d.num_circle_points = self.num_circle_points
        d.circle_radius = float(self.circle_radius)
        d.domain_extent = float(self.domain_extent)
        d.dt = float(self.dt)
        d.dealiasing_fraction = -1.0
        fr = getattr(self, "fourier_resident", None)
        d.flags = ((_lib.EXMI_FLAG_EXACT_FORM if getattr(self, "exact_form", False) else 0)
                   | (_lib.EXMI_FLAG_FOURIER_RESIDENT if fr is True else 0)
                   | (_lib.EXMI_FLAG_PHYSICAL_ROLLOUT if fr is False else 0))
        if self.order > 0:
            from .nonlin_fun import BaseNonlinearFun

            if not isinstance(self.nonlinear_fun, BaseNonlinearFun):
                raise NotImplementedError(
                    f"{type(self.nonlinear_fun).__name__}: only the library's nonlinear functions "
                    "(exponax_amd.nonlin_fun and the steppers' own) run in the fused kernels")
            self.nonlinear_fun.fill(d)
        return d

    def plan(self, batch: int) -> _lib.Plan:
        """The device plan for a batch of `batch` states (cached)."""
        if batch not in self._plans:
            self._plans[batch] = _lib.Plan(self._desc(batch), self.linear_operator)
        return self._plans[batch]

    # ------------------------------------------------------------- stepping
    @property
    def state_shape(self):
        return (self.num_channels,) + sp.spatial_shape(self.num_spatial_dims, self.num_points)

    @property
    def spectral_shape(self):
        return (self.num_channels,) + sp.wavenumber_shape(self.num_spatial_dims, self.num_points)

    def _split(self, t, trailing):
        if tuple(t.shape[-len(trailing):]) != tuple(trailing):
            raise ValueError(f"Expected shape (..., {', '.join(map(str, trailing))}), got {tuple(t.shape)}.")
        lead = tuple(t.shape[:-len(trailing)])
        return lead, int(np.prod(lead)) if lead else 1

    def step(self, u):
        """fft -> ETDRK step -> ifft (``_base_stepper.py:201-220``) on one
        state or a batch with leading axes."""
        import torch

        t, back = as_device_tensor(u)
        lead, B = self._split(t, self.state_shape)
        x = t.contiguous().to(torch.float32)
        out = torch.empty_like(x)
        n = int(np.prod(self.state_shape))
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_step(p.handle, x.data_ptr(), n, out.data_ptr(), n, _lib.stream_ptr()), "exmi_step")
        return restore_like(out, back)

    def step_fourier(self, u_hat):
        """``_base_stepper.py:222-239``."""
        import torch

        t, back = as_device_tensor(u_hat)
        lead, B = self._split(t, self.spectral_shape)
        x = t.contiguous().to(torch.complex64)
        out = torch.empty_like(x)
        n = int(np.prod(self.spectral_shape))
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_step_fourier(p.handle, x.data_ptr(), n, out.data_ptr(), n, _lib.stream_ptr()),
                   "exmi_step_fourier")
        return restore_like(out, back)

    def __call__(self, u):
        """``_base_stepper.py:241-271``: exactly one state."""
        shape = tuple(getattr(u, "shape", ()))
        if shape != self.state_shape:
            raise ValueError(f"Expected shape {self.state_shape}, got {shape}. For batched operation use "
                             "`exponax_amd.vmap` on this function.")
        return self.step(u)

    # ----------------------------------------------------- native rollouts
    def rollout_batch(self, u0, n: int, *, include_init: bool = False, use_graph: bool = True):
        """(..., C, N^D) -> (..., n[+1], C, N^D) in one captured graph."""
        import torch

        t, back = as_device_tensor(u0)
        lead, B = self._split(t, self.state_shape)
        x = t.contiguous().to(torch.float32)
        T = n + (1 if include_init else 0)
        out = torch.empty(lead + (T,) + self.state_shape, dtype=torch.float32, device=x.device)
        sz = int(np.prod(self.state_shape))
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_rollout(p.handle, x.data_ptr(), sz, out.data_ptr(), n, int(include_init),
                                           int(use_graph), _lib.stream_ptr()), "exmi_rollout")
        return restore_like(out, back)

    def rollout_checked(self, u0, n: int, *, bound: float = 0.0, include_init: bool = False):
        """``rollout_batch`` with failure detection: returns (trajectory,
        first_bad) where first_bad (B,) int32 holds, per trajectory, the first
        time index whose state has a non-finite value or one above `bound`
        in magnitude (bound <= 0: non-finite only), -1 if none
        (exmi_rollout_checked)."""
        import torch

        t, back = as_device_tensor(u0)
        lead, B = self._split(t, self.state_shape)
        x = t.contiguous().to(torch.float32)
        T = n + (1 if include_init else 0)
        out = torch.empty(lead + (T,) + self.state_shape, dtype=torch.float32, device=x.device)
        fb = torch.empty(B, dtype=torch.int32, device=x.device)
        sz = int(np.prod(self.state_shape))
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_rollout_checked(p.handle, x.data_ptr(), sz, out.data_ptr(), n, int(include_init),
                                                   float(bound), fb.data_ptr(), _lib.stream_ptr()),
                   "exmi_rollout_checked")
        return restore_like(out, back), restore_like(fb.reshape(lead) if lead else fb[0], back)

    def repeat_batch(self, u0, n: int, *, use_graph: bool = True):
        import torch

        t, back = as_device_tensor(u0)
        lead, B = self._split(t, self.state_shape)
        x = t.contiguous().to(torch.float32)
        out = torch.empty_like(x)
        sz = int(np.prod(self.state_shape))
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_repeat(p.handle, x.data_ptr(), sz, out.data_ptr(), n, int(use_graph),
                                          _lib.stream_ptr()), "exmi_repeat")
        return restore_like(out, back)

    # --------------------------------------------------------- derivatives
    def step_vjp(self, u, g):
        """(d step(u) / d u)^T g for one state or a batch (the reverse-mode
        derivative jax.vjp gives for ``_base_stepper.py:241-271``)."""
        import torch

        t, back = as_device_tensor(u)
        gt, _ = as_device_tensor(g)
        lead, B = self._split(t, self.state_shape)
        if tuple(gt.shape) != tuple(t.shape):
            raise ValueError(f"cotangent shape {tuple(gt.shape)} != state shape {tuple(t.shape)}")
        x = t.contiguous().to(torch.float32)
        gx = gt.contiguous().to(torch.float32)
        out = torch.empty_like(x)
        n = int(np.prod(self.state_shape))
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_step_vjp(p.handle, x.data_ptr(), n, gx.data_ptr(), n, out.data_ptr(), n,
                                            _lib.stream_ptr()), "exmi_step_vjp")
        return restore_like(out, back)

    def rollout_vjp(self, u0, n: int, *, g_traj=None, g_final=None, include_init: bool = False,
                    checkpoint_every: int = 10):
        """Gradient through ``rollout(n, include_init)`` (cotangents g_traj of
        the trajectory) and/or through ``repeat(n)`` (cotangent g_final of the
        last state), recomputing segments of `checkpoint_every` steps."""
        import torch

        t, back = as_device_tensor(u0)
        lead, B = self._split(t, self.state_shape)
        x = t.contiguous().to(torch.float32)
        sz = int(np.prod(self.state_shape))
        T = n + (1 if include_init else 0)
        gtp = gfp = None
        keep = []
        if g_traj is not None:
            gt, _ = as_device_tensor(g_traj)
            if tuple(gt.shape) != lead + (T,) + self.state_shape:
                raise ValueError(f"g_traj shape {tuple(gt.shape)} != {lead + (T,) + self.state_shape}")
            gt = gt.contiguous().to(torch.float32)
            keep.append(gt)
            gtp = gt.data_ptr()
        if g_final is not None:
            gf, _ = as_device_tensor(g_final)
            if tuple(gf.shape) != tuple(t.shape):
                raise ValueError(f"g_final shape {tuple(gf.shape)} != state shape {tuple(t.shape)}")
            gf = gf.contiguous().to(torch.float32)
            keep.append(gf)
            gfp = gf.data_ptr()
        out = torch.empty_like(x)
        p = self.plan(B)
        _lib.check(_lib.lib().exmi_rollout_vjp(p.handle, x.data_ptr(), sz, n, int(include_init), gtp, gfp,
                                               out.data_ptr(), int(checkpoint_every), _lib.stream_ptr()),
                   "exmi_rollout_vjp")
        return restore_like(out, back)

    def coefficient_tables(self, batch: int = 1) -> np.ndarray:
        """The plan's float32 (E, phi...) tables, see include/exmi.h."""
        return self.plan(batch).tables()


def _nonlinear_signature(stepper):
    """The nonlinear function's parameters as the plan carries them (one set
    per plan: the kernels take per-trajectory linear operators only)."""
    d = _lib.PlanDesc()
    stepper.nonlinear_fun.fill(d)
    return (type(stepper.nonlinear_fun), d.nonlin_kind, d.dealiasing_fraction, d.scale, d.conv_single_channel,
            d.conv_conservative, d.zero_mode_fix, d.num_poly, tuple(d.poly), tuple(d.general_scales),
            d.injection_mode, d.injection_scale, tuple(d.reaction_rates))


class Ensemble(BaseStepper):
    """A batch of steppers that differ only in their linear operator
    (per-trajectory coefficients), run in one launch per step — the
    ``eqx.filter_vmap`` over stepper constructors of exponax
    (docs/examples/performance_hints.ipynb cells 23-31)."""

    def __init__(self, members):
        members = list(members)
        if not members:
            raise ValueError("empty ensemble")
        m0 = members[0]
        for m in members[1:]:
            for attr in ("num_spatial_dims", "num_points", "num_channels", "dt", "order", "domain_extent",
                         "num_circle_points", "circle_radius"):
                if getattr(m, attr) != getattr(m0, attr):
                    raise ValueError(f"ensemble members differ in {attr}")
            if type(m.nonlinear_fun) is not type(m0.nonlinear_fun):
                raise ValueError("ensemble members differ in their nonlinear function")
            if _nonlinear_signature(m) != _nonlinear_signature(m0):
                raise NotImplementedError("ensemble members differ in their nonlinear function's parameters; "
                                          "per-trajectory coefficients are built for the linear operator only")
        self.__dict__.update({k: v for k, v in m0.__dict__.items() if k != "_plans"})
        self.members = members
        self.batch = len(members)
        lc = max(m.linear_operator.shape[0] for m in members)
        self.linear_operator = np.ascontiguousarray(
            np.stack([np.broadcast_to(m.linear_operator, (lc,) + m.linear_operator.shape[1:]) for m in members]))
        self._plans = {}

    def plan(self, batch: int) -> _lib.Plan:
        if batch != self.batch:
            raise ValueError(f"this ensemble advances exactly {self.batch} states, got a batch of {batch}")
        if batch not in self._plans:
            d = self._desc(batch, lin_batch=self.batch)
            d.lin_channels = self.linear_operator.shape[1]
            self._plans[batch] = _lib.Plan(d, self.linear_operator)
        return self._plans[batch]

    def __call__(self, u):
        return self.step(u)


def ensemble(make_stepper, params):
    """Build an :class:`Ensemble` from a constructor and per-member params,
    like ``eqx.filter_vmap(make_stepper)(params)``."""
    return Ensemble(make_stepper(p) for p in params)
