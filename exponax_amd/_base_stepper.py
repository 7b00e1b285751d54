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

            if not isinstance(getattr(self.nonlinear_fun, "_base", self.nonlinear_fun), BaseNonlinearFun):
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

    def fft_config(self, batch: int = 1):
        """Knobs of the runtime-size transforms of the plan for `batch`
        trajectories (grid sizes outside the fused kernels' set; None for the
        others): chosen by timing candidates when the plan is built."""
        return self.plan(batch).fft_config()


def _nonlinear_params(stepper):
    """The nonlinear function's parameters as the plan carries them: the ones
    ensemble members must share, and the vector of those the nonlinear term
    is linear in (its scales, polynomial coefficients and forcing)."""
    d = _lib.PlanDesc()
    stepper.nonlinear_fun.fill(d)
    fixed = (type(stepper.nonlinear_fun), d.nonlin_kind, d.dealiasing_fraction, d.conv_single_channel,
             d.conv_conservative, d.zero_mode_fix, d.num_poly, d.injection_mode, tuple(d.reaction_rates))
    vec = np.array([d.scale, *d.general_scales, *list(d.poly)[:d.num_poly], d.injection_scale], dtype=np.float64)
    return fixed, vec


def _common_factors(vecs):
    """f with vecs[b] = f[b] * vecs[0] for every member, or None."""
    v0 = vecs[0]
    j = int(np.argmax(np.abs(v0)))
    if v0[j] == 0.0:
        return np.ones(len(vecs)) if all(not np.any(v) for v in vecs) else None
    f = np.array([v[j] / v0[j] for v in vecs])
    for v, fb in zip(vecs, f):
        if not np.allclose(v, fb * v0, rtol=1e-12, atol=1e-300):
            return None
    return f


class _ScaledNonlinearFun:
    """The nonlinear function of an ensemble whose members' nonlinear terms
    differ by per-trajectory factors: member 0's function, its output
    multiplied by the factor of each trajectory of the batch."""

    def __init__(self, base, factors):
        self._base = base
        self.factors = np.asarray(factors, dtype=np.float64)

    def __getattr__(self, name):
        return getattr(self._base, name)

    def fill(self, desc):
        self._base.fill(desc)

    def __call__(self, u_hat):
        import torch

        t, back = as_device_tensor(u_hat)
        if t.dim() < 1 or t.shape[0] != len(self.factors):
            raise ValueError(f"this ensemble's nonlinear function takes a batch of {len(self.factors)} spectra")
        out, _ = as_device_tensor(self._base(t))
        f = torch.as_tensor(self.factors, dtype=torch.float32, device=out.device)
        return restore_like(out * f.reshape((-1,) + (1,) * (out.dim() - 1)), back)


class Ensemble(BaseStepper):
    """A batch of steppers with per-trajectory coefficients, run in one
    launch per step — the ``eqx.filter_vmap`` over stepper constructors of
    exponax (docs/examples/performance_hints.ipynb cells 23-31). Members may
    differ in their linear operator (per-trajectory operator tables) and in
    the parameters their nonlinear term is linear in, by a common factor per
    member (a convection or gradient-norm scale, all polynomial coefficients
    together, the Kolmogorov forcing with the convection scale: folded into
    the tables' nonlinear-term coefficients). Members that differ otherwise
    (e.g. one coefficient of several, or reaction rates) are advanced as
    groups of equal nonlinear parameters, one launch per group."""

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
        self.__dict__.update({k: v for k, v in m0.__dict__.items() if k != "_plans"})
        self.members = members
        self.batch = len(members)
        self._plans = {}
        self._groups = None
        self._factors = None
        if m0.order > 0:
            params = [_nonlinear_params(m) for m in members]
            same_fixed = all(p[0] == params[0][0] for p in params)
            f = _common_factors([p[1] for p in params]) if same_fixed else None
            if f is None:
                keys = {}
                for i, (fx, v) in enumerate(params):
                    keys.setdefault((fx, tuple(v)), []).append(i)
                self._groups = [(np.asarray(ix), Ensemble([members[i] for i in ix])) for ix in keys.values()]
            elif not np.all(f == 1.0):
                self._factors = f
                self.nonlinear_fun = _ScaledNonlinearFun(m0.nonlinear_fun, f)
        lc = max(m.linear_operator.shape[0] for m in members)
        self.linear_operator = np.ascontiguousarray(
            np.stack([np.broadcast_to(m.linear_operator, (lc,) + m.linear_operator.shape[1:]) for m in members]))

    def plan(self, batch: int) -> _lib.Plan:
        if self._groups is not None:
            raise NotImplementedError("this ensemble runs as groups of equal nonlinear parameters; "
                                      "it has one plan per group")
        if batch != self.batch:
            raise ValueError(f"this ensemble advances exactly {self.batch} states, got a batch of {batch}")
        if batch not in self._plans:
            d = self._desc(batch, lin_batch=self.batch)
            d.lin_channels = self.linear_operator.shape[1]
            self._plans[batch] = _lib.Plan(d, self.linear_operator, self._factors)
        return self._plans[batch]

    def _by_group(self, name, x, *args, **kw):
        """Run method `name` of every group on its members' states (the
        leading axis of x) and scatter the results back in member order."""
        import torch

        t, back = as_device_tensor(x)
        if t.dim() < 1 or t.shape[0] != self.batch:
            raise ValueError(f"this ensemble advances exactly {self.batch} states, got {tuple(t.shape)}")
        outs = None
        for ix, sub in self._groups:
            idx = torch.as_tensor(ix, device=t.device)
            r = getattr(sub, name)(t.index_select(0, idx), *args, **kw)
            rs = r if isinstance(r, tuple) else (r,)
            if outs is None:
                outs = [torch.empty((self.batch,) + tuple(v.shape[1:]), dtype=v.dtype, device=v.device) for v in rs]
            for o, v in zip(outs, rs):
                o.index_copy_(0, idx, v)
        res = tuple(restore_like(o, back) for o in outs)
        return res if len(res) > 1 else res[0]

    def step(self, u):
        return self._by_group("step", u) if self._groups is not None else super().step(u)

    def step_fourier(self, u_hat):
        return self._by_group("step_fourier", u_hat) if self._groups is not None else super().step_fourier(u_hat)

    def rollout_batch(self, u0, n: int, **kw):
        return self._by_group("rollout_batch", u0, n, **kw) if self._groups is not None \
            else super().rollout_batch(u0, n, **kw)

    def repeat_batch(self, u0, n: int, **kw):
        return self._by_group("repeat_batch", u0, n, **kw) if self._groups is not None \
            else super().repeat_batch(u0, n, **kw)

    def rollout_checked(self, u0, n: int, **kw):
        return self._by_group("rollout_checked", u0, n, **kw) if self._groups is not None \
            else super().rollout_checked(u0, n, **kw)

    def __call__(self, u):
        return self.step(u)


def ensemble(make_stepper, params):
    """Build an :class:`Ensemble` from a constructor and per-member params,
    like ``eqx.filter_vmap(make_stepper)(params)``."""
    return Ensemble(make_stepper(p) for p in params)
