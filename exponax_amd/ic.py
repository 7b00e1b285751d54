"""Seeded truncated-Fourier-series initial conditions (host, float64).

The reference draws its ICs from ``jax.random`` streams
(``ic/_truncated_fourier_series.py:17-100``). Those cannot be reproduced
without JAX, so the recipe here is the library's own:
- keep the modes with 1 <= max_d |k_d| <= cutoff from one half-plane, so that
  each cosine appears once;
- draw amplitudes a ~ N(0, 1) and phases phi ~ U[0, 2 pi);
- form u = sum a cos(2 pi k.x / L + phi), then divide by max |u|.
Trajectory b uses the b-th draws of one ``numpy.random.default_rng(seed)``
stream. The benchmark and the tests use this recipe.
"""
from __future__ import annotations

import itertools

import numpy as np


def _half_plane_modes(d, cutoff):
    modes = [k for k in itertools.product(range(-cutoff, cutoff + 1), repeat=d)
             if max(abs(c) for c in k) >= 1 and [c for c in reversed(k) if c != 0][0] > 0]
    return np.array(modes, dtype=np.float64)


def truncated_fourier_series(batch, num_spatial_dims, domain_extent, num_points, cutoff, rng):
    """(batch, 1, N, ..., N) float64 states; ``rng`` is a numpy Generator
    (or an int seed)."""
    if not isinstance(rng, np.random.Generator):
        rng = np.random.default_rng(rng)
    d = num_spatial_dims
    modes = _half_plane_modes(d, cutoff)
    g = np.linspace(0.0, domain_extent, num_points, endpoint=False)
    grids = np.meshgrid(*([g] * d), indexing="ij")
    phase_base = [sum(modes[m, i] * grids[i] for i in range(d)) * (2 * np.pi / domain_extent)
                  for m in range(len(modes))]
    out = np.empty((batch, 1) + (num_points,) * d)
    for b in range(batch):
        a = rng.standard_normal(len(modes))
        ph = rng.uniform(0.0, 2 * np.pi, len(modes))
        u = np.zeros((num_points,) * d)
        for m in range(len(modes)):
            u += a[m] * np.cos(phase_base[m] + ph[m])
        out[b, 0] = u / np.max(np.abs(u))
    return out


def truncated_fourier_series_device(batch, num_spatial_dims, domain_extent, num_points, cutoff, seed, *,
                                    num_channels=1, device="cuda"):
    """The same recipe on the GPU: (batch, C, N, ..., N) float32 torch tensor.

    The amplitudes and phases are drawn on the host from the same
    ``numpy.random.default_rng(seed)`` stream and in the same order as the
    host recipe (per trajectory, per channel: amplitudes, then phases); the
    device writes the spectra, inverse-transforms them and normalises each
    field by its max |u| (``exmi_ic_tfs``). ``domain_extent`` does not enter:
    on the grid x = i L / N the phases are 2 pi k . i / N."""
    import torch

    from . import _lib

    if cutoff < 1 or cutoff >= num_points // 2:
        raise ValueError(f"cutoff must be in 1..{num_points // 2 - 1}, got {cutoff}")
    rng = seed if isinstance(seed, np.random.Generator) else np.random.default_rng(seed)
    modes = np.ascontiguousarray(_half_plane_modes(num_spatial_dims, cutoff).astype(np.int32))
    nm, nf = len(modes), batch * num_channels
    amp = np.empty((nf, nm), dtype=np.float32)
    phase = np.empty((nf, nm), dtype=np.float32)
    for f in range(nf):
        amp[f] = rng.standard_normal(nm)
        phase[f] = rng.uniform(0.0, 2 * np.pi, nm)
    a = torch.tensor(amp, device=device)
    p = torch.tensor(phase, device=device)
    out = torch.empty((batch, num_channels) + (num_points,) * num_spatial_dims, dtype=torch.float32, device=device)
    _lib.check(_lib.lib().exmi_ic_tfs(num_spatial_dims, num_points, nf, nm, modes.ctypes.data, a.data_ptr(),
                                      p.data_ptr(), out.data_ptr(), _lib.stream_ptr()), "exmi_ic_tfs")
    return out
