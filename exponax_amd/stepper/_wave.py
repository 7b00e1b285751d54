"""Wave equation stepper (``stepper/_wave.py:9-197``): h_t = v, v_t = c^2
Laplace h, solved exactly per Fourier mode through the reference's
diagonalisation (``exmi_wave_step_fourier``), in 1-D, 2-D and 3-D."""
from __future__ import annotations

from .. import _lib
from .. import spectral as sp
from .._tensor import as_device_tensor, restore_like


class Wave:
    def __init__(self, num_spatial_dims: int, domain_extent: float, num_points: int, dt: float, *,
                 speed_of_sound: float = 1.0):
        self.num_spatial_dims = num_spatial_dims
        self.domain_extent = domain_extent
        self.num_points = num_points
        self.dt = dt
        self.dx = domain_extent / num_points
        self.num_channels = 2
        self.order = 0
        self.speed_of_sound = speed_of_sound

    @property
    def state_shape(self):
        return (2,) + sp.spatial_shape(self.num_spatial_dims, self.num_points)

    @property
    def spectral_shape(self):
        return (2,) + sp.wavenumber_shape(self.num_spatial_dims, self.num_points)

    def step_fourier(self, u_hat):
        """``_wave.py:175-197`` on (..., 2, Nr) spectra."""
        import torch

        t, back = as_device_tensor(u_hat)
        if tuple(t.shape[t.dim() - len(self.spectral_shape):]) != self.spectral_shape:
            raise ValueError(f"Expected shape (..., {self.spectral_shape}), got {tuple(t.shape)}")
        x = t.to(torch.complex64).contiguous().clone()
        batch = x.numel() // (2 * int(__import__("numpy").prod(self.spectral_shape[1:])))
        _lib.check(_lib.lib().exmi_wave_step_fourier(self.num_spatial_dims, self.num_points, batch,
                                                     float(self.domain_extent), float(self.speed_of_sound),
                                                     float(self.dt), x.data_ptr(), _lib.stream_ptr()),
                   "exmi_wave_step_fourier")
        return restore_like(x, back)

    def step(self, u):
        """One step of a state or a batch with leading axes."""
        import torch

        t, back = as_device_tensor(u)
        D = self.num_spatial_dims
        uh = sp.fft(t.to(torch.float32), num_spatial_dims=D)
        return restore_like(sp.ifft(self.step_fourier(uh), num_spatial_dims=D, num_points=self.num_points), back)

    def __call__(self, u):
        shape = tuple(getattr(u, "shape", ()))
        if shape != self.state_shape:
            raise ValueError(f"Expected shape {self.state_shape}, got {shape}. For batched operation use "
                             "`exponax_amd.vmap` on this function.")
        return self.step(u)

    def repeat_batch(self, u0, n: int):
        u = u0
        for _ in range(n):
            u = self.step(u)
        return u

    def rollout_batch(self, u0, n: int, *, include_init: bool = False):
        import torch

        t, back = as_device_tensor(u0)
        k = t.dim() - len(self.state_shape)
        out = [t.to(torch.float32)] if include_init else []
        u = t
        for _ in range(n):
            u = self.step(u)
            out.append(u)
        return restore_like(torch.stack(out, dim=k), back)


__all__ = ["Wave"]
