"""CPU restatement of exponax's spectral utilities (test infrastructure only).

Every function follows a function of ``exponax/_spectral.py`` (exponax 0.2.0);
the citation is given per function. Arrays are NumPy; ``dtype`` selects the
working precision (float64 by default, float32 for the FP32 oracle).
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

_WORKERS = 1


def set_fft_workers(n: int) -> None:
    """Thread count for scipy.fft (pocketfft) in the oracle."""
    global _WORKERS
    _WORKERS = int(n)


def complex_of(dtype) -> np.dtype:
    return np.dtype(np.complex64) if np.dtype(dtype) == np.float32 else np.dtype(np.complex128)


def build_wavenumbers(num_spatial_dims: int, num_points: int) -> np.ndarray:
    """Integer wavenumber grid, shape (D, N, ..., N//2+1).

    Follows ``_spectral.py:13-49``: last axis ``rfftfreq`` (0..N//2), other
    axes ``fftfreq`` (0..N/2-1, -N/2..-1), stacked with ``meshgrid(indexing="ij")``.
    """
    last = np.fft.rfftfreq(num_points, 1.0 / num_points)
    other = np.fft.fftfreq(num_points, 1.0 / num_points)
    axes = [other] * (num_spatial_dims - 1) + [last]
    return np.stack(np.meshgrid(*axes, indexing="ij"))


def build_scaled_wavenumbers(num_spatial_dims, domain_extent, num_points):
    """``_spectral.py:52-83``: wavenumbers times 2*pi/L."""
    return (2 * np.pi / domain_extent) * build_wavenumbers(num_spatial_dims, num_points)


def build_derivative_operator(num_spatial_dims, domain_extent, num_points):
    """``_spectral.py:86-115``: i*kappa_d, Nyquist entries kept."""
    return 1j * build_scaled_wavenumbers(num_spatial_dims, domain_extent, num_points)


def build_laplace_operator(derivative_operator, *, order: int = 2):
    """``_spectral.py:118-155``: sum_d (i kappa_d)^order, order even; order 0 -> ones."""
    if order % 2 != 0:
        raise ValueError("Order must be even.")
    if order == 0:
        return np.ones((1,) + derivative_operator.shape[1:], dtype=derivative_operator.dtype)
    return np.sum(derivative_operator**order, axis=0, keepdims=True)


def build_gradient_inner_product_operator(derivative_operator, velocity, *, order: int = 1):
    """``_spectral.py:158-209``: sum_i velocity_i (i kappa_i)^order, order odd."""
    if order % 2 != 1:
        raise ValueError("Order must be odd.")
    velocity = np.asarray(velocity)
    if velocity.shape != (derivative_operator.shape[0],):
        raise ValueError(
            f"Expected velocity shape to be {derivative_operator.shape[0]}, got {velocity.shape}."
        )
    return np.einsum("i,i...->...", velocity, derivative_operator**order)[None, ...]


def spatial_shape(num_spatial_dims, num_points):
    """``_spectral.py:231-251``."""
    return (num_points,) * num_spatial_dims


def wavenumber_shape(num_spatial_dims, num_points):
    """``_spectral.py:254-275``."""
    return (num_points,) * (num_spatial_dims - 1) + (num_points // 2 + 1,)


def low_pass_filter_mask(num_spatial_dims, num_points, *, cutoff, axis_separate=True):
    """``_spectral.py:278-342``: keep |k_d| <= cutoff on every axis (hypercube),
    or |k|_2 <= cutoff when ``axis_separate`` is False. Shape (1, ..., N//2+1)."""
    wn = build_wavenumbers(num_spatial_dims, num_points)
    if axis_separate:
        mask = np.ones(wn.shape[1:], dtype=bool)
        for grid in wn:
            mask &= np.abs(grid) <= cutoff
    else:
        mask = np.linalg.norm(wn, axis=0) <= cutoff
    return mask[None, ...]


def oddball_filter_mask(num_spatial_dims, num_points):
    """``_spectral.py:345-400``: zero the Nyquist modes on even grids."""
    if num_points % 2 == 1:
        return np.ones((1,) + wavenumber_shape(num_spatial_dims, num_points), dtype=bool)
    return low_pass_filter_mask(num_spatial_dims, num_points, cutoff=num_points // 2 - 1)


def dealiasing_mask(num_spatial_dims, num_points, dealiasing_fraction):
    """Dealias mask of ``nonlin_fun/_base.py:59-71``.

    cutoff = fraction * (N//2) - 1 (a float), kept modes |k_d| <= cutoff.
    ``None`` fraction means no mask.
    """
    if dealiasing_fraction is None:
        return None
    highest_resolved = num_points // 2
    start_aliased = dealiasing_fraction * highest_resolved
    return low_pass_filter_mask(num_spatial_dims, num_points, cutoff=start_aliased - 1)


def _scaling_array(num_spatial_dims, num_points, right_den, other_den):
    """``_spectral.py:403-466``."""
    last = np.fft.rfftfreq(num_points, 1.0 / num_points)
    other = np.fft.fftfreq(num_points, 1.0 / num_points)
    s_last = np.where(last == 0, num_points, num_points / right_den)
    s_other = np.where(other == 0, num_points, num_points / other_den)
    if num_points % 2 == 0:
        s_last = np.where(last == num_points // 2, num_points, s_last)
        s_other = np.where(other == -num_points // 2, num_points, s_other)
    axes = [s_other] * (num_spatial_dims - 1) + [s_last]
    return np.prod(np.stack(np.meshgrid(*axes, indexing="ij")), axis=0, keepdims=True)


def build_scaling_array(num_spatial_dims, num_points, *, mode):
    """``_spectral.py:469-527``."""
    if mode == "norm_compensation":
        return _scaling_array(num_spatial_dims, num_points, 1, 1)
    if mode == "reconstruction":
        return _scaling_array(num_spatial_dims, num_points, 2, 1)
    if mode == "coef_extraction":
        return _scaling_array(num_spatial_dims, num_points, 2, 2)
    raise ValueError("Invalid mode.")


def get_modes_slices(num_spatial_dims, num_points):
    """``_spectral.py:530-611``: slice blocks of the retained modes."""
    from itertools import product

    even = num_points % 2 == 0
    nyq = num_points // 2
    if even:
        left, right = slice(None, nyq), slice(-nyq, None)
    else:
        left, right = slice(None, nyq + 1), slice(-nyq, None)
    blocks = [[slice(None, nyq + 1)]] + [[left, right] for _ in range(num_spatial_dims - 1)]
    return tuple(tuple([slice(None)] + list(reversed(p))) for p in product(*blocks))


def _axes(num_spatial_dims):
    return tuple(range(-num_spatial_dims, 0))


def fft(field, *, num_spatial_dims=None):
    """``_spectral.py:614-656``: rfftn over the last D axes, norm="backward"."""
    if num_spatial_dims is None:
        num_spatial_dims = field.ndim - 1
    return sfft.rfftn(field, axes=_axes(num_spatial_dims), workers=_WORKERS)


def ifft(field_hat, *, num_spatial_dims=None, num_points=None):
    """``_spectral.py:659-721``: irfftn with s=(N,)*D, 1/N^D scaling.

    The c2r convention (Im of last-axis DC/Nyquist ignored) is pocketfft's,
    the same as numpy.fft.irfftn and jnp.fft.irfftn.
    """
    if num_spatial_dims is None:
        num_spatial_dims = field_hat.ndim - 1
    if num_points is None:
        if num_spatial_dims >= 2:
            num_points = field_hat.shape[-2]
        else:
            raise ValueError("num_points must be provided if num_spatial_dims == 1.")
    return sfft.irfftn(
        field_hat,
        s=spatial_shape(num_spatial_dims, num_points),
        axes=_axes(num_spatial_dims),
        workers=_WORKERS,
    )


def derivative(field, domain_extent, *, order=1):
    """``_spectral.py:724-788``: spectral derivative of every channel."""
    channels = field.shape[0]
    d = field.ndim - 1
    n = field.shape[1]
    op = build_derivative_operator(d, domain_extent, n) ** order
    fh = fft(field, num_spatial_dims=d)
    if channels == 1:
        dh = op * fh
    else:
        dh = fh[:, None] * op[None, ...]
    return ifft(dh, num_spatial_dims=d, num_points=n)


def get_spectrum(state, *, power=True, radial_binning="sum"):
    """``_spectral.py:866-1030``: (power) spectrum with radial shell binning."""
    d = state.ndim - 1
    n = state.shape[-1]
    mag_abs = np.abs(fft(state, num_spatial_dims=d))
    magnitude = mag_abs / build_scaling_array(d, n, mode="reconstruction")
    if power:
        comp = mag_abs / build_scaling_array(d, n, mode="norm_compensation")
        quantity = 0.5 * magnitude * comp
    else:
        quantity = magnitude
    if d == 1:
        return quantity
    wn = build_wavenumbers(d, n)
    wn1 = build_wavenumbers(1, n)[0]
    knorm = np.linalg.norm(wn, axis=0)
    dk = wn1[1] - wn1[0]
    out = np.zeros((state.shape[0], wn1.shape[0]), dtype=quantity.dtype)
    for i, k in enumerate(wn1):
        m = (knorm >= k - dk / 2) & (knorm < k + dk / 2)
        for c in range(state.shape[0]):
            sel = quantity[c][m]
            if radial_binning == "average":
                out[c, i] = np.nanmean(sel) if sel.size else np.nan
            else:
                out[c, i] = np.nansum(sel)
    return out


def map_between_resolutions(state, new_num_points, *, oddball_zero=True):
    """``_interpolation.py:195-285``: spectral up/down-sampling."""
    d = state.ndim - 1
    old_n = state.shape[-1]
    c = state.shape[0]
    if old_n == new_num_points:
        return state
    old_hat = fft(state, num_spatial_dims=d) / build_scaling_array(d, old_n, mode="norm_compensation")
    if new_num_points > old_n and old_n % 2 == 0 and oddball_zero:
        old_hat = old_hat * oddball_filter_mask(d, old_n)
    new_hat = np.zeros((c,) + wavenumber_shape(d, new_num_points), dtype=old_hat.dtype)
    for block in get_modes_slices(d, min(old_n, new_num_points)):
        new_hat[block] = old_hat[block]
    new_hat = new_hat * build_scaling_array(d, new_num_points, mode="norm_compensation")
    if old_n > new_num_points and new_num_points % 2 == 0 and oddball_zero:
        new_hat = new_hat * oddball_filter_mask(d, new_num_points)
    return ifft(new_hat, num_spatial_dims=d, num_points=new_num_points)


def make_grid(num_spatial_dims, domain_extent, num_points, *, full=False, zero_centered=False):
    """``_utils.py:10-60``: (D, N, ..., N) coordinates, right end excluded."""
    if full:
        g = np.linspace(0, domain_extent, num_points + 1, endpoint=True)
    else:
        g = np.linspace(0, domain_extent, num_points, endpoint=False)
    if zero_centered:
        g = g - domain_extent / 2
    return np.stack(np.meshgrid(*([g] * num_spatial_dims), indexing="ij"))
