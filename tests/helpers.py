"""Shared helpers for the parity tests (oracle construction, tolerances)."""
import numpy as np

from oracle import ic as oic
from oracle import stepper as ost

# rel-L2 gate of the FP32 GPU path against the FP32 oracle (one and ten steps)
RTOL_STEP = 1e-5


def rel_l2(a, b):
    """Sum over channels of ||a - b|| / ||b|| (metrics/_spatial.py:537-586),
    averaged over a leading batch axis when present."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == b.ndim and a.ndim >= 3:
        return float(np.mean([ost.nrmse(x, y) for x, y in zip(a, b)]))
    return ost.nrmse(a, b)


def max_rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max([ost.nrmse(x, y) for x, y in zip(a, b)]))


def ics(batch, D, L, N, cutoff, seed, channels=1):
    u0, rng = oic.tfs_batch(batch, D, L, N, cutoff, seed, num_channels=channels)
    return u0.astype(np.float32), rng
