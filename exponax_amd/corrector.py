"""Learned corrector for a coarse, large-dt stepper.

The network is a periodic 1-D CNN: y = x + f(x), where f is 5 Conv1d layers
of kernel 5 (1 -> 32 -> 32 -> 32 -> 32 -> 1), with GELU (tanh form) after
layers 1-4. It runs in the library's BF16 matrix-core kernels
(``exmi_cnn_forward`` / ``exmi_cnn_backward``).

A :class:`CorrectedStepper` applies it after every step of a coarse stepper:
u_{t+1} = cnn(coarse(u_t)). It is trained on trajectories that the GPU fine
stepper generates, downsampled with ``map_between_resolutions``
(``_interpolation.py:195-285``).

The reference has no learned component. The nearest reference code is the
periodic Conv1d emulator of
docs/examples/learning_burgers_autoregressive_neural_operator.ipynb.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._tensor import as_device_tensor, restore_like

NUM_PARAMS = _lib.EXMI_CNN_NUM_PARAMS
CH, KW = 32, 5
_SHAPES = [((CH, 1, KW), (CH,)), ((CH, CH, KW), (CH,)), ((CH, CH, KW), (CH,)), ((CH, CH, KW), (CH,)),
           ((1, CH, KW), (1,))]


def init_params(seed: int = 0, scale_last: float = 0.1) -> np.ndarray:
    """He-normal weights, zero biases; the last layer scaled so that the
    untrained network is a small correction."""
    rng = np.random.default_rng(seed)
    ps = []
    for i, (ws, bs) in enumerate(_SHAPES):
        w = rng.standard_normal(ws) * np.sqrt(2.0 / (ws[1] * ws[2]))
        if i == len(_SHAPES) - 1:
            w = w * scale_last
        ps += [w.ravel(), np.zeros(int(np.prod(bs)))]
    return np.concatenate(ps).astype(np.float32)


def _check_n(n):
    if n not in (32, 64, 128):
        raise NotImplementedError(f"the corrector network is built for 32, 64 or 128 points, got {n}")


def cnn_forward(params: torch.Tensor, x: torch.Tensor, *, workspace: torch.Tensor | None = None) -> torch.Tensor:
    """y = x + f(x) for x of shape (..., N) (float32, on the device)."""
    _check_n(x.shape[-1])
    x = x.contiguous()
    y = torch.empty_like(x)
    count = x.numel() // x.shape[-1]
    _lib.check(_lib.lib().exmi_cnn_forward(params.data_ptr(), x.shape[-1], count, x.data_ptr(), y.data_ptr(),
                                           None if workspace is None else workspace.data_ptr(), _lib.stream_ptr()),
               "exmi_cnn_forward")
    return y


def workspace_for(n: int, count: int, device) -> torch.Tensor:
    nbytes = int(_lib.lib().exmi_cnn_workspace_bytes(n, count))
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


class _CnnFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params):
        n = x.shape[-1]
        count = x.numel() // n
        ws = workspace_for(n, count, x.device)
        y = cnn_forward(params, x, workspace=ws)
        ctx.save_for_backward(params)
        ctx.ws, ctx.shape = ws, x.shape
        return y

    @staticmethod
    def backward(ctx, dy):
        (params,) = ctx.saved_tensors
        dy = dy.contiguous()
        n = ctx.shape[-1]
        count = dy.numel() // n
        dx = torch.empty_like(dy) if ctx.needs_input_grad[0] else None
        dp = torch.empty_like(params)
        _lib.check(_lib.lib().exmi_cnn_backward(params.data_ptr(), n, count, dy.data_ptr(),
                                                None if dx is None else dx.data_ptr(), dp.data_ptr(),
                                                ctx.ws.data_ptr(), _lib.stream_ptr()), "exmi_cnn_backward")
        ctx.ws = None
        return dx, dp


class PeriodicCNN:
    """The corrector network; ``params`` is a float32 device tensor of
    ``NUM_PARAMS`` values in the PyTorch Conv1d layout (see include/exmi.h)."""

    def __init__(self, params=None, *, seed: int = 0, device="cuda"):
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (NUM_PARAMS,):
            raise ValueError(f"expected {NUM_PARAMS} parameters, got {p.shape}")
        self.params = torch.tensor(p, device=device, requires_grad=True)

    def __call__(self, x):
        """Inference: y = x + f(x) on (..., N) states (any array kind)."""
        t, back = as_device_tensor(x)
        with torch.no_grad():
            y = cnn_forward(self.params.detach(), t.to(torch.float32))
        return restore_like(y, back)

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """Differentiable in x and in ``params`` (torch autograd)."""
        return _CnnFn.apply(x.contiguous(), self.params)


class CorrectedStepper:
    """u_{t+1} = cnn(coarse(u_t)) for a 1-D single-channel coarse stepper."""

    def __init__(self, coarse, cnn: PeriodicCNN):
        if coarse.num_spatial_dims != 1 or coarse.num_channels != 1:
            raise NotImplementedError("the corrector is built for 1-D single-channel steppers")
        _check_n(coarse.num_points)
        self.coarse = coarse
        self.cnn = cnn

    def step(self, u):
        t, back = as_device_tensor(u)
        return restore_like(self.cnn(self.coarse.step(t)), back)

    __call__ = step

    def rollout_batch(self, u0, n: int, *, include_init: bool = False):
        """(B, 1, N) -> (B, n[+1], 1, N) through exmi_corrected_rollout."""
        t, back = as_device_tensor(u0)
        u = t.contiguous().to(torch.float32)
        lead, B = self.coarse._split(u, self.coarse.state_shape)
        N = self.coarse.num_points
        T = n + (1 if include_init else 0)
        out = torch.empty(lead + (T, 1, N), dtype=torch.float32, device=u.device)
        p = self.coarse.plan(B)
        _lib.check(_lib.lib().exmi_corrected_rollout(p.handle, self.cnn.params.data_ptr(), u.data_ptr(), N,
                                                     out.data_ptr(), n, int(include_init), _lib.stream_ptr()),
                   "exmi_corrected_rollout")
        return restore_like(out, back)


# ------------------------------------------------------------------ training
def downsampled_trajectories(fine, u0, n_coarse: int, stride: int, coarse_points: int) -> torch.Tensor:
    """Fine rollout of ``n_coarse * stride`` steps from u0 (B, 1, Nf); every
    ``stride``-th state (the initial one included) mapped to ``coarse_points``
    with ``map_between_resolutions``. Returns (B, n_coarse + 1, 1, Nc) on the
    device."""
    from .spectral import map_between_resolutions

    traj = fine.rollout_batch(u0, n_coarse * stride, include_init=True)
    sel = traj[:, ::stride].contiguous()
    return map_between_resolutions(sel, coarse_points)


def train(cnn: PeriodicCNN, coarse, data: torch.Tensor, *, iters: int = 2000, batch: int = 512, unroll: int = 4,
          lr: float = 1e-3, clip: float = 1.0, seed: int = 0, log_every: int = 0):
    """Fit ``cnn`` so that u_{t+j} = cnn(coarse(u_{t+j-1})) follows the data
    for ``unroll`` steps. Windows are drawn at random from
    ``data`` (B, T+1, 1, N). The loss is the mean squared error over the
    unrolled steps. Gradients flow through the network (exmi_cnn_backward) and
    through the coarse solver (exmi_step_vjp). Adam, cosine learning-rate
    decay, gradient-norm clipping. Returns the loss history."""
    from . import autograd

    B, T1 = data.shape[0], data.shape[1]
    if T1 <= unroll:
        raise ValueError("trajectories shorter than the unroll length")
    opt = torch.optim.Adam([cnn.params], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=iters)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for it in range(iters):
        b = torch.randint(0, B, (batch,), generator=gen).to(data.device)
        t0 = torch.randint(0, T1 - unroll, (batch,), generator=gen).to(data.device)
        u = data[b, t0].contiguous()
        loss = 0.0
        for j in range(1, unroll + 1):
            u = cnn.apply(autograd.step(coarse, u))
            loss = loss + ((u - data[b, t0 + j]) ** 2).mean()
        loss = loss / unroll
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_([cnn.params], clip)
        opt.step()
        sched.step()
        if log_every and (it % log_every == 0 or it == iters - 1):
            hist.append((it, float(loss.item())))
    return hist


def rollout_with_fallback(corrected: CorrectedStepper, fine, u0_fine, n_coarse: int, stride: int,
                          *, bound: float = 10.0, u0_coarse=None):
    """Corrected coarse rollout of (B, 1, Nf) initial states. Trajectories
    that leave [-bound, bound] or become non-finite are recomputed by the
    plain fine stepper and downsampled, so the fine solver stays the fallback
    and the reference. ``u0_coarse`` (default: u0_fine mapped to the coarse grid)
    are the coarse initial states. Returns (trajectory (B, n_coarse, 1, Nc),
    mask of the trajectories that fell back)."""
    from .spectral import map_between_resolutions

    from ._base_stepper import Ensemble
    from .metrics import first_bad_step

    nc = corrected.coarse.num_points
    u0c = map_between_resolutions(u0_fine, nc) if u0_coarse is None else u0_coarse
    tr = corrected.rollout_batch(u0c, n_coarse)
    bad = first_bad_step(tr, bound) >= 0  # exmi_first_bad_step on the device
    if bool(bad.any()):
        idx = torch.nonzero(bad).flatten()
        sub = fine
        if isinstance(fine, Ensemble):  # the failed trajectories' own members (per-trajectory coefficients)
            sub = Ensemble([fine.members[i] for i in idx.tolist()])
        ft = downsampled_trajectories(sub, u0_fine[idx].contiguous(), n_coarse, stride, nc)
        tr[idx] = ft[:, 1:]
    return tr, bad
