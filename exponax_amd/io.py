"""On-disk trajectory datasets.

A dataset is the float32 array ``(B, T, C, N, ..., N)`` that a vmapped
``rollout`` produces (exponax/_utils.py:92-186, trajectory axis after the
batch axis), stored as a NumPy ``.npy`` file, plus a JSON sidecar
``<path>.json`` describing the stepper and the sampling. Files are written
chunk by chunk from device rollouts, so a dataset larger than device or host
memory streams through a memory map; readers get a read-only memory map.
Continuing a rollout from its last stored state reproduces the continuous
rollout exactly (the multi-step kernels do the same arithmetic per step).
"""
from __future__ import annotations

import json
import os

import numpy as np


def _stepper_meta(stepper):
    keys = ("num_spatial_dims", "domain_extent", "num_points", "num_channels", "dt", "order", "num_sub_steps")
    meta = {k: getattr(stepper, k) for k in keys if hasattr(stepper, k)}
    meta["stepper"] = type(stepper).__name__
    inner = getattr(stepper, "stepper", None)
    if inner is not None:
        meta["sub_stepper"] = type(inner).__name__
    return meta


def rollout_to_file(stepper, u0, n: int, path: str, *, include_init: bool = True, chunk: int = 256,
                    meta: dict | None = None) -> str:
    """Roll `stepper` (an exponax_amd stepper, ensemble or RepeatedStepper)
    out from the batch `u0` (B, C, N, ..., N) for `n` steps and store the
    trajectories at `path` (``.npy``), `chunk` steps per device rollout.
    Returns the path of the JSON sidecar."""
    import torch

    from ._tensor import as_device_tensor

    if n < 0 or chunk < 1:
        raise ValueError("n must be >= 0 and chunk >= 1")
    x, _ = as_device_tensor(u0)
    x = x.contiguous().to(torch.float32)
    state_shape = tuple(stepper.state_shape)
    if tuple(x.shape[1:]) != state_shape:
        raise ValueError(f"Expected u0 of shape (B, {', '.join(map(str, state_shape))}), got {tuple(x.shape)}.")
    B = x.shape[0]
    T = n + (1 if include_init else 0)
    arr = np.lib.format.open_memmap(path, mode="w+", dtype=np.float32, shape=(B, T) + state_shape)
    t = 0
    if include_init:
        arr[:, 0] = x.cpu().numpy()
        t = 1
    u = x
    done = 0
    while done < n:
        m = min(chunk, n - done)
        tr = stepper.rollout_batch(u, m)
        arr[:, t:t + m] = tr.cpu().numpy()
        u = tr[:, m - 1].contiguous()
        t += m
        done += m
    arr.flush()
    del arr
    side = dict(_stepper_meta(stepper))
    side.update({"batch": int(B), "steps": int(n), "include_init": bool(include_init),
                 "shape": [int(B), int(T)] + [int(s) for s in state_shape], "dtype": "float32",
                 "layout": "(batch, time, channel, space...)"})
    if meta:
        side["user"] = meta
    spath = path + ".json"
    with open(spath, "w") as f:
        json.dump(side, f, indent=1)
    return spath


def open_trajectories(path: str):
    """(read-only memory map of the (B, T, C, N, ..., N) array, sidecar dict
    or None)."""
    arr = np.load(path, mmap_mode="r", allow_pickle=False)
    side = None
    if os.path.exists(path + ".json"):
        with open(path + ".json") as f:
            side = json.load(f)
        if side.get("shape") and list(arr.shape) != side["shape"]:
            raise ValueError(f"{path}: array shape {arr.shape} does not match its sidecar {side['shape']}")
    return arr, side


__all__ = ["open_trajectories", "rollout_to_file"]
