"""Nonlinear-term descriptors with the constructor signatures of exponax's
``nonlin_fun/`` classes.

A descriptor records the kind and scalars of N(u); the fused step kernels
evaluate it. Calling a descriptor on a spectrum evaluates N(u_hat) on the
device through ``exmi_nonlinear`` (``BaseNonlinearFun.__call__``,
``nonlin_fun/_base.py:139-159``).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import spectral as sp
from ._tensor import as_device_tensor, restore_like


def domain_extent_of(derivative_operator, num_spatial_dims, num_points) -> float:
    """Recover L from a standard derivative operator 1j*2*pi/L*k; custom
    operators are not supported by the fused kernels (they compute i*kappa
    from the mode index)."""
    d = np.asarray(derivative_operator)
    if d.shape != (num_spatial_dims,) + sp.wavenumber_shape(num_spatial_dims, num_points):
        raise ValueError(f"derivative_operator has shape {d.shape}")
    k1 = d[(num_spatial_dims - 1,) + (0,) * (num_spatial_dims - 1) + (1,)]
    if not np.imag(k1) > 0:
        raise NotImplementedError("only the standard spectral derivative operator is supported")
    L = 2 * np.pi / float(np.imag(k1))
    ref = sp.build_derivative_operator(num_spatial_dims, L, num_points)
    if not np.allclose(d, ref, rtol=1e-6, atol=1e-6 * np.max(np.abs(ref))):
        raise NotImplementedError("only the standard spectral derivative operator is supported")
    return L


class BaseNonlinearFun:
    """``nonlin_fun/_base.py:9-159``."""

    kind = _lib.EXMI_NL_ZERO

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction=None):
        self.num_spatial_dims = num_spatial_dims
        self.num_points = num_points
        self.dealiasing_fraction = dealiasing_fraction
        if dealiasing_fraction is None:
            self.dealiasing_mask = None
        else:
            cutoff = dealiasing_fraction * (num_points // 2) - 1
            self.dealiasing_mask = sp.low_pass_filter_mask(num_spatial_dims, num_points, cutoff=cutoff)
        self.domain_extent = 1.0
        self._plans = {}

    def fill(self, desc: _lib.PlanDesc) -> None:
        desc.nonlin_kind = self.kind
        desc.dealiasing_fraction = -1.0 if self.dealiasing_fraction is None else float(self.dealiasing_fraction)

    def dealias(self, u_hat):
        """``_base.py:73-97``."""
        if self.dealiasing_mask is None:
            raise ValueError("Nonlinear function was set up without dealiasing")
        import torch

        t, back = as_device_tensor(u_hat)
        m = torch.as_tensor(self.dealiasing_mask, device=t.device)
        return restore_like(t * m, back)

    def __call__(self, u_hat):
        """Evaluate N(u_hat) for (C, Nr) or batched (..., C, Nr) spectra."""
        import torch

        t, back = as_device_tensor(u_hat)
        D, N = self.num_spatial_dims, self.num_points
        ws = sp.wavenumber_shape(D, N)
        if tuple(t.shape[-D:]) != ws:
            raise ValueError(f"Expected spectral shape (..., C, {ws}), got {tuple(t.shape)}")
        C = t.shape[-D - 1]
        lead = tuple(t.shape[:-D - 1])
        B = int(np.prod(lead)) if lead else 1
        plan = self._plan(C, B)
        x = t.contiguous().to(torch.complex64)
        out = torch.empty_like(x)
        nr = int(np.prod(ws))
        _lib.check(_lib.lib().exmi_nonlinear(plan.handle, x.data_ptr(), C * nr, out.data_ptr(), C * nr,
                                             _lib.stream_ptr()), "exmi_nonlinear")
        return restore_like(out, back)

    def _plan(self, C, B):
        key = (C, B)
        if key not in self._plans:
            d = _lib.PlanDesc()
            d.num_spatial_dims = self.num_spatial_dims
            d.num_points = self.num_points
            d.num_channels = C
            d.batch = B
            d.order = 1
            d.lin_batch = 1
            d.lin_channels = 1
            d.num_circle_points = 1
            d.circle_radius = 1.0
            d.domain_extent = self.domain_extent
            d.dt = 1.0
            self.fill(d)
            nr = int(np.prod(sp.wavenumber_shape(self.num_spatial_dims, self.num_points)))
            self._plans[key] = _lib.Plan(d, np.zeros(nr))
        return self._plans[key]


class ZeroNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_zero.py:7-38``."""

    kind = _lib.EXMI_NL_ZERO

    def __init__(self, num_spatial_dims, num_points):
        super().__init__(num_spatial_dims, num_points)


class ConvectionNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_convection.py:7-245``."""

    kind = _lib.EXMI_NL_CONVECTION

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction=2 / 3, scale=1.0,
                 single_channel=False, conservative=False):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.domain_extent = domain_extent_of(derivative_operator, num_spatial_dims, num_points)
        self.scale = scale
        self.single_channel = single_channel
        self.conservative = conservative

    def fill(self, d):
        super().fill(d)
        d.scale = float(self.scale)
        d.conv_single_channel = int(bool(self.single_channel))
        d.conv_conservative = int(bool(self.conservative))


class GradientNormNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_gradient_norm.py:8-101``."""

    kind = _lib.EXMI_NL_GRADIENT_NORM

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction, zero_mode_fix=True,
                 scale=1.0):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.domain_extent = domain_extent_of(derivative_operator, num_spatial_dims, num_points)
        self.zero_mode_fix = zero_mode_fix
        self.scale = scale

    def fill(self, d):
        super().fill(d)
        d.scale = float(self.scale)
        d.zero_mode_fix = int(bool(self.zero_mode_fix))


class PolynomialNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_polynomial.py:6-76``."""

    kind = _lib.EXMI_NL_POLYNOMIAL

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction, coefficients):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.coefficients = tuple(float(c) for c in coefficients)
        if len(self.coefficients) > _lib.EXMI_MAX_POLY:
            raise NotImplementedError(f"at most {_lib.EXMI_MAX_POLY} polynomial coefficients")

    def fill(self, d):
        super().fill(d)
        d.num_poly = len(self.coefficients)
        for i, c in enumerate(self.coefficients):
            d.poly[i] = c


class GeneralNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_general_nonlinear.py:9-119``."""

    kind = _lib.EXMI_NL_GENERAL

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction,
                 scale_list=(0.0, -1.0, 0.0), zero_mode_fix=True):
        if len(scale_list) != 3:
            raise ValueError("The scale list must have exactly 3 elements")
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.domain_extent = domain_extent_of(derivative_operator, num_spatial_dims, num_points)
        self.scale_list = tuple(float(s) for s in scale_list)
        self.zero_mode_fix = zero_mode_fix

    def fill(self, d):
        super().fill(d)
        for i in range(3):
            d.general_scales[i] = self.scale_list[i]
        d.zero_mode_fix = int(bool(self.zero_mode_fix))


class VorticityConvection2d(BaseNonlinearFun):
    """``nonlin_fun/_vorticity_convection.py:8-99``."""

    kind = _lib.EXMI_NL_VORTICITY

    def __init__(self, num_spatial_dims, num_points, *, convection_scale=1.0, derivative_operator,
                 dealiasing_fraction):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}.")
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.domain_extent = domain_extent_of(derivative_operator, num_spatial_dims, num_points)
        self.convection_scale = convection_scale

    def fill(self, d):
        super().fill(d)
        d.scale = float(self.convection_scale)


class VorticityConvection2dKolmogorov(VorticityConvection2d):
    """``nonlin_fun/_vorticity_convection.py:102-182``."""

    kind = _lib.EXMI_NL_KOLMOGOROV

    def __init__(self, num_spatial_dims, num_points, *, convection_scale=1.0, injection_mode=4, injection_scale=1.0,
                 derivative_operator, dealiasing_fraction):
        super().__init__(num_spatial_dims, num_points, convection_scale=convection_scale,
                         derivative_operator=derivative_operator, dealiasing_fraction=dealiasing_fraction)
        self.injection_mode = injection_mode
        self.injection_scale = injection_scale

    def fill(self, d):
        super().fill(d)
        d.injection_mode = int(self.injection_mode)
        d.injection_scale = float(self.injection_scale)


class ProjectedConvection3d(BaseNonlinearFun):
    """``nonlin_fun/_projected_convection.py:19-136``: P(u x omega) with the
    Leray projection P (``_leray.py``), three channels, 3-D only."""

    kind = _lib.EXMI_NL_PROJECTED_CONVECTION

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction=2 / 3):
        if num_spatial_dims != 3:
            raise ValueError("ProjectedConvection3d only supports 3 spatial dimensions.")
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.domain_extent = domain_extent_of(derivative_operator, num_spatial_dims, num_points)


class ProjectedConvection3dKolmogorov(ProjectedConvection3d):
    """``nonlin_fun/_projected_convection.py:139-226``: plus the forcing of
    intensity injection_scale at wavenumber (0, injection_mode, 0) of the
    first velocity channel."""

    kind = _lib.EXMI_NL_PROJECTED_CONVECTION_KOLMOGOROV

    def __init__(self, num_spatial_dims, num_points, *, injection_mode=4, injection_scale=1.0, derivative_operator,
                 dealiasing_fraction):
        super().__init__(num_spatial_dims, num_points, derivative_operator=derivative_operator,
                         dealiasing_fraction=dealiasing_fraction)
        self.injection_mode = injection_mode
        self.injection_scale = injection_scale

    def fill(self, d):
        super().fill(d)
        d.injection_mode = int(self.injection_mode)
        d.injection_scale = float(self.injection_scale)
