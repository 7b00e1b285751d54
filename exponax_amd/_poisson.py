"""``Poisson`` (``_poisson.py:12-80``): -Laplace^(order/2) u = f solved in
Fourier space with the mean-zero convention, on the GPU
(``exmi_poisson_fourier``)."""
from __future__ import annotations

from . import _lib
from . import spectral as sp
from ._tensor import as_device_tensor, restore_like


class Poisson:
    def __init__(self, num_spatial_dims: int, domain_extent: float, num_points: int, *, order: int = 2):
        if order < 2 or order % 2:
            raise ValueError(f"order must be a positive even integer, got {order}")
        self.num_spatial_dims = num_spatial_dims
        self.domain_extent = domain_extent
        self.num_points = num_points
        self.dx = domain_extent / num_points
        self.order = order

    def step_fourier(self, f_hat):
        """``_poisson.py:44-49``: -f_hat / Laplace (0 at the mean mode)."""
        import numpy as np
        import torch

        t, back = as_device_tensor(f_hat)
        ws = sp.wavenumber_shape(self.num_spatial_dims, self.num_points)
        if tuple(t.shape[t.dim() - len(ws):]) != ws:
            raise ValueError(f"Expected spectra (..., {ws}), got {tuple(t.shape)}")
        x = t.to(torch.complex64).contiguous().clone()
        count = x.numel() // int(np.prod(ws))
        _lib.check(_lib.lib().exmi_poisson_fourier(self.num_spatial_dims, self.num_points, count,
                                                   float(self.domain_extent), int(self.order), x.data_ptr(),
                                                   _lib.stream_ptr()), "exmi_poisson_fourier")
        return restore_like(x, back)

    def step(self, f):
        """``_poisson.py:51-63``."""
        import torch

        t, back = as_device_tensor(f)
        D = self.num_spatial_dims
        fh = sp.fft(t.to(torch.float32), num_spatial_dims=D)
        return restore_like(sp.ifft(self.step_fourier(fh), num_spatial_dims=D, num_points=self.num_points), back)

    def __call__(self, f):
        """``_poisson.py:65-80``: a state (C, N, ..., N)."""
        shape = tuple(getattr(f, "shape", ()))
        if shape[1:] != sp.spatial_shape(self.num_spatial_dims, self.num_points):
            raise ValueError(f"Shape of f[1:] is {shape[1:]} but should be "
                             f"{sp.spatial_shape(self.num_spatial_dims, self.num_points)}")
        return self.step(f)


__all__ = ["Poisson"]
