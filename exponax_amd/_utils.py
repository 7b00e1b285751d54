"""Rollout / repeat / vmap with exponax's semantics (``_utils.py:92-254``).

For exponax_amd steppers (single or vmapped) the whole trajectory is one
captured hipGraph of per-step launches (``exmi_rollout``); any other
callable is looped on the host.
"""
from __future__ import annotations

import numpy as np

from ._base_stepper import BaseStepper, Ensemble
from ._repeated_stepper import RepeatedStepper
from .stepper._wave import Wave

_STEPPERS = (BaseStepper, RepeatedStepper, Wave)


class vmap:
    """Batch a stepper over a leading axis: ``vmap(stepper)(u_batch)``
    advances all states in one launch (the reference's ``jax.vmap``)."""

    def __init__(self, stepper):
        self.stepper = stepper

    def __call__(self, u):
        if isinstance(self.stepper, _STEPPERS):
            return self.stepper.step(u)
        return _stack([self.stepper(x) for x in u])


def _stack(xs):
    import torch

    if isinstance(xs[0], torch.Tensor):
        return torch.stack(xs)
    return np.stack(xs)


def _native(stepper_fn):
    if isinstance(stepper_fn, vmap) and isinstance(stepper_fn.stepper, _STEPPERS):
        return stepper_fn.stepper, True
    if isinstance(stepper_fn, Ensemble):
        return stepper_fn, True
    if isinstance(stepper_fn, _STEPPERS):
        return stepper_fn, False
    return None, False


def _time_first(traj, lead_ndim):
    """(lead..., T, ...) -> (T, lead..., ...): the stacking of ``lax.scan``
    over a batched carry."""
    return traj.movedim(lead_ndim, 0) if hasattr(traj, "movedim") else np.moveaxis(traj, lead_ndim, 0)


def rollout(stepper_fn, n: int, *, include_init: bool = False, takes_aux: bool = False, constant_aux: bool = True):
    """``_utils.py:92-186``. Returns a function u0 -> trajectory (n[+1], ...):
    time is the leading axis, also for a batched stepper (``vmap(stepper)``
    or an :class:`Ensemble`), whose trajectory is (n[+1], B, ...) as the
    reference's ``rollout(jax.vmap(stepper), n)`` scans over the batched
    state. ``stepper.rollout_batch`` keeps the batch axis first
    ((B, n[+1], ...), the layout of ``jax.vmap(rollout(stepper, n))``)."""
    if takes_aux:
        def run_aux(u0, aux):
            out = [u0] if include_init else []
            u = u0
            for i in range(n):
                a = aux if constant_aux else _index(aux, i)
                u = stepper_fn(u, a)
                out.append(u)
            return _stack(out)
        return run_aux

    st, batched = _native(stepper_fn)

    def run(u0):
        if st is not None:
            if not batched:
                shape = tuple(getattr(u0, "shape", ()))
                if shape != st.state_shape:
                    raise ValueError(f"Expected shape {st.state_shape}, got {shape}.")
                return st.rollout_batch(u0, n, include_init=include_init)
            lead_ndim = len(getattr(u0, "shape", ())) - len(st.state_shape)
            return _time_first(st.rollout_batch(u0, n, include_init=include_init), lead_ndim)
        out = [u0] if include_init else []
        u = u0
        for _ in range(n):
            u = stepper_fn(u)
            out.append(u)
        return _stack(out)
    return run


def repeat(stepper_fn, n: int, *, takes_aux: bool = False, constant_aux: bool = True):
    """``_utils.py:189-254``. Returns a function u0 -> state after n steps."""
    if takes_aux:
        def run_aux(u0, aux):
            u = u0
            for i in range(n):
                u = stepper_fn(u, aux if constant_aux else _index(aux, i))
            return u
        return run_aux

    st, batched = _native(stepper_fn)

    def run(u0):
        if st is not None:
            if not batched:
                shape = tuple(getattr(u0, "shape", ()))
                if shape != st.state_shape:
                    raise ValueError(f"Expected shape {st.state_shape}, got {shape}.")
            return st.repeat_batch(u0, n)
        u = u0
        for _ in range(n):
            u = stepper_fn(u)
        return u
    return run


def _index(aux, i):
    if isinstance(aux, (tuple, list)):
        return type(aux)(_index(a, i) for a in aux)
    if isinstance(aux, dict):
        return {k: _index(v, i) for k, v in aux.items()}
    return aux[i]


def make_grid(num_spatial_dims, domain_extent, num_points, *, full=False, zero_centered=False):
    """``_utils.py:10-60`` (host NumPy, float64)."""
    g = (np.linspace(0, domain_extent, num_points + 1, endpoint=True) if full
         else np.linspace(0, domain_extent, num_points, endpoint=False))
    if zero_centered:
        g = g - domain_extent / 2
    return np.stack(np.meshgrid(*([g] * num_spatial_dims), indexing="ij"))


def rollout_group(steppers, u0s, n: int, *, include_init: bool = False, use_graph=True, out=None):
    """Roll several steppers (e.g. ensembles on different grid sizes)
    together. ``u0s[i]`` is a (B_i, C, N_i^D) batch for ``steppers[i]``;
    returns the list of (B_i, n[+1], C, N_i^D) trajectories (written into
    ``out`` when given). ``use_graph``: False/0 = in order on the current
    stream; True/1 = one captured graph with a parallel branch per stepper;
    2 = one cached graph per stepper, launched concurrently on the steppers'
    own streams; 3 = as 2 without graphs (include/exmi.h)."""
    import ctypes

    import torch

    from . import _lib
    from ._tensor import as_device_tensor

    plans, xs, outs, bs = [], [], [], []
    for st, u0 in zip(steppers, u0s):
        t, _ = as_device_tensor(u0)
        lead, B = st._split(t, st.state_shape)
        x = t.contiguous().to(torch.float32)
        T = n + (1 if include_init else 0)
        shape = lead + (T,) + st.state_shape
        if out is not None:
            o = out[len(outs)]
            if tuple(o.shape) != shape or o.dtype != torch.float32 or not o.is_contiguous():
                raise ValueError(f"out[{len(outs)}] must be a contiguous float32 tensor of shape {shape}")
            outs.append(o)
        else:
            outs.append(torch.empty(shape, dtype=torch.float32, device=x.device))
        xs.append(x)
        plans.append(st.plan(B))
        bs.append(int(np.prod(st.state_shape)))
    k = len(plans)
    P = (ctypes.c_void_p * k)(*[p.handle.value for p in plans])
    U = (ctypes.c_void_p * k)(*[x.data_ptr() for x in xs])
    S = (ctypes.c_int64 * k)(*bs)
    O = (ctypes.c_void_p * k)(*[o.data_ptr() for o in outs])
    _lib.check(_lib.lib().exmi_rollout_group(P, k, U, S, O, n, int(include_init), int(use_graph),
                                             _lib.stream_ptr()), "exmi_rollout_group")
    return outs
