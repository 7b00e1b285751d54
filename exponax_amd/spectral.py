"""Spectral grids, operators and transforms.

Setup-time helpers (wavenumbers, derivative/Laplace operators, masks,
scaling arrays) run once per stepper on the host in float64 NumPy, with the
signatures of exponax's ``_spectral.py`` so that user steppers'
``_build_linear_operator`` code ports unchanged. ``fft``/``ifft`` run the
gfx950 transform kernels on device tensors.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._tensor import as_device_tensor, restore_like


def build_wavenumbers(num_spatial_dims: int, num_points: int) -> np.ndarray:
    """(D, N, ..., N//2+1) integer grid; ``_spectral.py:13-49``."""
    last = np.fft.rfftfreq(num_points, 1.0 / num_points)
    other = np.fft.fftfreq(num_points, 1.0 / num_points)
    return np.stack(np.meshgrid(*([other] * (num_spatial_dims - 1) + [last]), indexing="ij"))


def build_scaled_wavenumbers(num_spatial_dims, domain_extent, num_points):
    """``_spectral.py:52-83``."""
    return (2 * np.pi / domain_extent) * build_wavenumbers(num_spatial_dims, num_points)


def build_derivative_operator(num_spatial_dims, domain_extent, num_points):
    """i * kappa, shape (D, ..., N//2+1); ``_spectral.py:86-115``."""
    return 1j * build_scaled_wavenumbers(num_spatial_dims, domain_extent, num_points)


def build_laplace_operator(derivative_operator, *, order: int = 2):
    """sum_d (i kappa_d)^order; ``_spectral.py:118-155``."""
    if order % 2 != 0:
        raise ValueError("Order must be even.")
    if order == 0:
        return np.ones((1,) + derivative_operator.shape[1:], dtype=derivative_operator.dtype)
    return np.sum(derivative_operator**order, axis=0, keepdims=True)


def build_gradient_inner_product_operator(derivative_operator, velocity, *, order: int = 1):
    """``_spectral.py:158-209``."""
    if order % 2 != 1:
        raise ValueError("Order must be odd.")
    velocity = np.asarray(velocity)
    if velocity.shape != (derivative_operator.shape[0],):
        raise ValueError(f"Expected velocity shape to be {derivative_operator.shape[0]}, got {velocity.shape}.")
    return np.tensordot(velocity, derivative_operator**order, axes=(0, 0))[None]


def space_indices(num_spatial_dims):
    return tuple(range(-num_spatial_dims, 0))


def spatial_shape(num_spatial_dims, num_points):
    return (num_points,) * num_spatial_dims


def wavenumber_shape(num_spatial_dims, num_points):
    return (num_points,) * (num_spatial_dims - 1) + (num_points // 2 + 1,)


def low_pass_filter_mask(num_spatial_dims, num_points, *, cutoff, axis_separate=True):
    """``_spectral.py:278-342``."""
    wn = build_wavenumbers(num_spatial_dims, num_points)
    if axis_separate:
        mask = np.all(np.abs(wn) <= cutoff, axis=0)
    else:
        mask = np.linalg.norm(wn, axis=0) <= cutoff
    return mask[None]


def oddball_filter_mask(num_spatial_dims, num_points):
    """``_spectral.py:345-400``."""
    if num_points % 2 == 1:
        return np.ones((1,) + wavenumber_shape(num_spatial_dims, num_points), dtype=bool)
    return low_pass_filter_mask(num_spatial_dims, num_points, cutoff=num_points // 2 - 1)


def build_scaling_array(num_spatial_dims, num_points, *, mode):
    """``_spectral.py:403-527``."""
    dens = {"norm_compensation": (1, 1), "reconstruction": (2, 1), "coef_extraction": (2, 2)}
    if mode not in dens:
        raise ValueError("Invalid mode.")
    rden, oden = dens[mode]
    last = np.fft.rfftfreq(num_points, 1.0 / num_points)
    other = np.fft.fftfreq(num_points, 1.0 / num_points)
    sl = np.where(last == 0, num_points, num_points / rden)
    so = np.where(other == 0, num_points, num_points / oden)
    if num_points % 2 == 0:
        sl = np.where(last == num_points // 2, num_points, sl)
        so = np.where(other == -num_points // 2, num_points, so)
    grids = np.meshgrid(*([so] * (num_spatial_dims - 1) + [sl]), indexing="ij")
    return np.prod(np.stack(grids), axis=0, keepdims=True)


def fft(field, *, num_spatial_dims=None):
    """rfftn over the last D axes, norm="backward"; ``_spectral.py:614-656``.

    Runs the gfx950 real-FFT kernel. Accepts torch CUDA tensors (returned as
    complex64 tensors), NumPy arrays or DLPack producers (returned in kind).
    """
    t, back = as_device_tensor(field)
    if num_spatial_dims is None:
        num_spatial_dims = t.dim() - 1
    if num_spatial_dims not in (1, 2, 3):
        raise NotImplementedError("fft: num_spatial_dims 1, 2 and 3 are built")
    import torch

    n = t.shape[-1]
    if any(s != n for s in t.shape[t.dim() - num_spatial_dims:]):
        raise ValueError(f"expected N points on each of the last {num_spatial_dims} axes, got {tuple(t.shape)}")
    x = t.contiguous().to(torch.float32)
    count = x.numel() // n ** num_spatial_dims
    out = torch.empty(x.shape[:-1] + (n // 2 + 1,), dtype=torch.complex64, device=x.device)
    _lib.check(_lib.lib().exmi_rfft(num_spatial_dims, n, count, x.data_ptr(), out.data_ptr(), _lib.stream_ptr()),
               "exmi_rfft")
    return restore_like(out, back)


def ifft(field_hat, *, num_spatial_dims=None, num_points=None):
    """irfftn with s=(N,)*D; ``_spectral.py:659-721``."""
    t, back = as_device_tensor(field_hat)
    if num_spatial_dims is None:
        num_spatial_dims = t.dim() - 1
    if num_points is None:
        if num_spatial_dims >= 2:
            num_points = t.shape[-2]
        else:
            raise ValueError("num_points must be provided if num_spatial_dims == 1.")
    if num_spatial_dims not in (1, 2, 3):
        raise NotImplementedError("ifft: num_spatial_dims 1, 2 and 3 are built")
    import torch

    if t.shape[-1] != num_points // 2 + 1:
        raise ValueError(f"expected {num_points // 2 + 1} modes on the last axis, got {t.shape[-1]}")
    if any(s != num_points for s in t.shape[t.dim() - num_spatial_dims:-1]):
        raise ValueError(f"expected {num_points} modes on the leading spectral axes, got {tuple(t.shape)}")
    z = t.contiguous().to(torch.complex64)
    count = z.numel() // (z.shape[-1] * num_points ** (num_spatial_dims - 1))
    out = torch.empty(z.shape[:-1] + (num_points,), dtype=torch.float32, device=z.device)
    _lib.check(_lib.lib().exmi_irfft(num_spatial_dims, num_points, count, z.data_ptr(), out.data_ptr(),
                                     _lib.stream_ptr()), "exmi_irfft")
    return restore_like(out, back)


def _spatial_dims(shape) -> int:
    """Trailing axes of the last axis' length, at most len(shape) - 1: the
    reference's D = ndim - 1 for a single state (C, N, ..., N), and the same D
    for leading batch axes of another length."""
    d = 0
    while d < len(shape) - 1 and shape[len(shape) - 1 - d] == shape[-1]:
        d += 1
    return max(d, 1)


def get_spectrum(state, *, power: bool = True, radial_binning: str = "sum", num_spatial_dims=None):
    """(Power) spectrum of a state ``(C, N, ..., N)`` -> ``(C, N//2+1)``;
    ``_spectral.py:866-1030``. In 2-D and 3-D the modes are binned by the
    norm of their wavenumber ("sum" or "average"). Leading batch axes are
    allowed; ``num_spatial_dims`` defaults to the number of trailing axes of
    length N, at most ``state.ndim - 1`` (the reference's single state)."""
    t, back = as_device_tensor(state)
    if t.dim() < 2:
        raise ValueError("expected a state with a leading channel axis")
    if radial_binning not in ("sum", "average"):
        raise ValueError(f"Invalid radial_binning {radial_binning!r}")
    import torch

    D = _spatial_dims(t.shape) if num_spatial_dims is None else int(num_spatial_dims)
    n = t.shape[-1]
    if any(s != n for s in t.shape[t.dim() - D:]):
        raise ValueError(f"expected N points on each of the last {D} axes, got {tuple(t.shape)}")
    x = t.contiguous().to(torch.float32)
    count = x.numel() // n**D
    out = torch.empty(x.shape[:t.dim() - D] + (n // 2 + 1,), dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().exmi_spectrum_binned(D, n, count, x.data_ptr(), out.data_ptr(), int(power),
                                               int(radial_binning == "average"), _lib.stream_ptr()), "exmi_spectrum")
    return restore_like(out, back)


def map_between_resolutions(state, new_num_points: int, *, oddball_zero: bool = True, num_spatial_dims=None):
    """Spectral up/down-sampling of a state ``(C, N, ..., N)`` to
    ``new_num_points`` per axis; ``_interpolation.py:195-285``. Leading batch
    axes are allowed; ``num_spatial_dims`` defaults as in ``get_spectrum``."""
    t, back = as_device_tensor(state)
    if t.dim() < 2:
        raise ValueError("expected a state with a leading channel axis")
    import torch

    D = _spatial_dims(t.shape) if num_spatial_dims is None else int(num_spatial_dims)
    n = t.shape[-1]
    if any(s != n for s in t.shape[t.dim() - D:]):
        raise ValueError(f"expected N points on each of the last {D} axes, got {tuple(t.shape)}")
    x = t.contiguous().to(torch.float32)
    count = x.numel() // n**D
    out = torch.empty(x.shape[:t.dim() - D] + (new_num_points,) * D, dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().exmi_map_resolution(D, n, new_num_points, count, x.data_ptr(), out.data_ptr(),
                                              int(oddball_zero), _lib.stream_ptr()), "exmi_map_resolution")
    return restore_like(out, back)
