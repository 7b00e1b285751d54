"""Seeded initial conditions for the parity and benchmark workloads.

The reference draws its ICs from ``jax.random`` streams
(``exponax/ic/_truncated_fourier_series.py:17-100``), which cannot be
reproduced without JAX; these recipes are the project's own, defined here
once and used by the tests, the fixtures and the benchmark.

``tfs(num_spatial_dims, domain_extent, num_points, cutoff, rng)``: truncated
Fourier series. Modes with 1 <= max_d |k_d| <= cutoff, taken from one
half-plane (k_last > 0, or k_last == 0 and the first non-zero leading
component > 0) so that each cosine appears once; amplitude ~ N(0, 1), phase
~ U[0, 2 pi); u = sum a cos(2 pi k.x / L + phi) (zero mean by construction);
then u /= max |u|. Drawn in float64, returned float64.
"""
from __future__ import annotations

import itertools

import numpy as np


def _half_plane_modes(num_spatial_dims, cutoff):
    rng1 = range(-cutoff, cutoff + 1)
    modes = []
    for k in itertools.product(rng1, repeat=num_spatial_dims):
        if max(abs(c) for c in k) < 1:
            continue
        # canonical representative of {k, -k}: first non-zero component from
        # the last axis backwards is positive
        nz = [c for c in reversed(k) if c != 0]
        if nz[0] > 0:
            modes.append(k)
    return np.array(modes, dtype=np.float64)


def tfs(num_spatial_dims, domain_extent, num_points, cutoff, rng, *, num_channels=1):
    """One state (C, N, ..., N) from the truncated-Fourier-series recipe."""
    modes = _half_plane_modes(num_spatial_dims, cutoff)
    g = np.linspace(0.0, domain_extent, num_points, endpoint=False)
    grids = np.meshgrid(*([g] * num_spatial_dims), indexing="ij")
    out = np.empty((num_channels,) + (num_points,) * num_spatial_dims)
    for c in range(num_channels):
        a = rng.standard_normal(len(modes))
        phi = rng.uniform(0.0, 2 * np.pi, len(modes))
        u = np.zeros((num_points,) * num_spatial_dims)
        for m in range(len(modes)):
            arg = sum(modes[m, d] * grids[d] for d in range(num_spatial_dims))
            u += a[m] * np.cos(2 * np.pi * arg / domain_extent + phi[m])
        out[c] = u / np.max(np.abs(u))
    return out


def tfs_batch(batch, num_spatial_dims, domain_extent, num_points, cutoff, seed, *, num_channels=1):
    """(B, C, N, ..., N) ICs; trajectory b uses the b-th draws of one stream."""
    rng = np.random.default_rng(seed)
    return np.stack([tfs(num_spatial_dims, domain_extent, num_points, cutoff, rng, num_channels=num_channels)
                     for _ in range(batch)]), rng
