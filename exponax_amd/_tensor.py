"""Array interop: torch CUDA tensors are native; NumPy arrays and DLPack
producers (JAX arrays) are moved to the device and results are handed back
in the caller's kind."""
from __future__ import annotations

import numpy as np


def as_device_tensor(x):
    """Return (CUDA tensor, restore-info). Raises ExmiError without a HIP
    device: the kernels are the only compute path."""
    import torch

    from . import _lib

    _lib.lib()

    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            return x, ("torch", None)
        return x.cuda(), ("torch_cpu", None)
    if isinstance(x, np.ndarray) or np.isscalar(x):
        return torch.as_tensor(np.asarray(x)).cuda(), ("numpy", None)
    if hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
        if not t.is_cuda:
            t = t.cuda()
        return t, ("dlpack", type(x))
    return torch.as_tensor(np.asarray(x)).cuda(), ("numpy", None)


def restore_like(t, info):
    kind, _ = info
    if kind == "torch":
        return t
    if kind == "torch_cpu":
        return t.cpu()
    if kind == "numpy":
        return t.cpu().numpy()
    if kind == "dlpack":
        try:
            import jax.dlpack  # noqa: F401

            return jax.dlpack.from_dlpack(t)
        except Exception:
            return t
    return t
