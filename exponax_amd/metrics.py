"""Error metrics on the GPU: ``exponax.metrics.nRMSE``
(``metrics/_spatial.py:537-586``) for one state pair, and batched over
leading axes (the ``jax.vmap`` of it; ``mean_metric`` is its mean); and the
first failing state of every trajectory of a batch (failure detection)."""
from __future__ import annotations

import numpy as np

from . import _lib
from ._tensor import as_device_tensor, restore_like


def nRMSE_batch(u_pred, u_ref):
    """Per-sample nRMSE of ``(..., C, N^D)`` pairs with the channel axis at
    ``-(D+1)``: sum over channels of ||u - u_ref|| / ||u_ref||. The spatial
    rank D is taken as the trailing axes after the channel axis, so callers
    pass states shaped (B, C, N) or (B, C, N, N)."""
    import torch

    p, back = as_device_tensor(u_pred)
    r, _ = as_device_tensor(u_ref)
    if tuple(p.shape) != tuple(r.shape):
        raise ValueError(f"shape mismatch {tuple(p.shape)} vs {tuple(r.shape)}")
    if p.dim() < 3:
        raise ValueError("expected (B, C, N...) states")
    B, C = p.shape[0], p.shape[1]
    n = int(np.prod(p.shape[2:]))
    pc = p.contiguous().to(torch.float32)
    rc = r.contiguous().to(torch.float32)
    out = torch.empty(B, dtype=torch.float32, device=pc.device)
    _lib.check(_lib.lib().exmi_nrmse(B, C, n, pc.data_ptr(), C * n, rc.data_ptr(), C * n, out.data_ptr(),
                                     _lib.stream_ptr()), "exmi_nrmse")
    return restore_like(out, back)


def nRMSE(u_pred, u_ref, *, domain_extent: float = 1.0) -> float:
    """One (C, N^D) pair; ``domain_extent`` cancels in the ratio."""
    p, _ = as_device_tensor(u_pred)
    r, _ = as_device_tensor(u_ref)
    return float(nRMSE_batch(p[None], r[None])[0].item())


def mean_nRMSE(u_pred, u_ref) -> float:
    """``metrics.mean_metric(nRMSE, ...)``: the batch mean."""
    return float(nRMSE_batch(u_pred, u_ref).mean().item())


def first_bad_step(traj, bound: float = 0.0):
    """Per trajectory of ``(B, T, ...)`` states: the first time index whose
    state holds a value that is not finite or exceeds ``bound`` in magnitude
    (``bound <= 0``: non-finite values only), -1 if none; (B,) int32 on the
    device (exmi_first_bad_step)."""
    import torch

    t, back = as_device_tensor(traj)
    if t.dim() < 3:
        raise ValueError("expected (B, T, ...) trajectories")
    x = t.contiguous().to(torch.float32)
    B, T = x.shape[0], x.shape[1]
    n = int(np.prod(x.shape[2:]))
    out = torch.empty(B, dtype=torch.int32, device=x.device)
    _lib.check(_lib.lib().exmi_first_bad_step(B, T, n, x.data_ptr(), T * n, n, float(bound), out.data_ptr(),
                                              _lib.stream_ptr()), "exmi_first_bad_step")
    return restore_like(out, back)
