"""Generic stepper families (``stepper/generic/_linear.py``, ``_convection.py``,
``_gradient_norm.py``, ``_polynomial.py``, ``_nonlinear.py``,
``_vorticity_convection.py``): a linear part sum_j a_j sum_d (i kappa_d)^j
with one of the library's nonlinear functions; the Normalized* variants fix
L = 1, dt = 1 and the Difficulty* variants derive the normalized values from
difficulties (``_utils.py``). Odd j make L complex (1-D kernels)."""
from __future__ import annotations

import warnings

import numpy as np

from ..._base_stepper import BaseStepper
from ...nonlin_fun import (ConvectionNonlinearFun, GeneralNonlinearFun, GradientNormNonlinearFun,
                           PolynomialNonlinearFun, VorticityConvection2d, VorticityConvection2dKolmogorov,
                           ZeroNonlinearFun)
from . import _utils as u


def _linear_from_coefficients(derivative_operator, coefficients):
    # sum_j a_j sum_d (i kappa_d)^j, keeping the channel axis (e.g. _linear.py)
    out = np.zeros((1,) + derivative_operator.shape[1:], dtype=np.complex128)
    for j, a in enumerate(coefficients):
        out = out + np.sum(a * derivative_operator**j, axis=0, keepdims=True)
    return out


class _Coefficients(BaseStepper):
    def _build_linear_operator(self, derivative_operator):
        return _linear_from_coefficients(derivative_operator, self.linear_coefficients)


# ------------------------------------------------------------------ linear
class GeneralLinearStepper(_Coefficients):
    """``_linear.py``: ETDRK0, zero nonlinearity."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, linear_coefficients=(0.0, -0.1, 0.01)):
        self.linear_coefficients = tuple(linear_coefficients)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=0)

    def _build_nonlinear_fun(self, derivative_operator):
        return ZeroNonlinearFun(self.num_spatial_dims, self.num_points)


class NormalizedLinearStepper(GeneralLinearStepper):
    def __init__(self, num_spatial_dims, num_points, *, normalized_linear_coefficients=(0.0, -0.5, 0.01)):
        self.normalized_linear_coefficients = tuple(normalized_linear_coefficients)
        super().__init__(num_spatial_dims, 1.0, num_points, 1.0, linear_coefficients=normalized_linear_coefficients)


class DifficultyLinearStepper(NormalizedLinearStepper):
    def __init__(self, num_spatial_dims=1, num_points=48, *, linear_difficulties=(0.0, -2.0)):
        self.linear_difficulties = tuple(linear_difficulties)
        super().__init__(num_spatial_dims, num_points, normalized_linear_coefficients=(
            u.extract_normalized_coefficients_from_difficulty(linear_difficulties, num_spatial_dims=num_spatial_dims,
                                                              num_points=num_points)))


class DifficultyLinearStepperSimple(DifficultyLinearStepper):
    """One derivative of order `order` with difficulty `difficulty`."""

    def __init__(self, num_spatial_dims=1, num_points=48, *, difficulty=-2.0, order=1):
        super().__init__(num_spatial_dims, num_points, linear_difficulties=(0.0,) * order + (difficulty,))


def DiffultyLinearStepperSimple(*args, **kwargs):  # noqa: N802 - the reference's deprecated alias
    warnings.warn("`DiffultyLinearStepperSimple` is deprecated due to a typo. Use `DifficultyLinearStepperSimple` "
                  "instead.", DeprecationWarning, stacklevel=2)
    return DifficultyLinearStepperSimple(*args, **kwargs)


# -------------------------------------------------------------- convection
class GeneralConvectionStepper(_Coefficients):
    """``_convection.py:12-152``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, linear_coefficients=(0.0, 0.0, 0.01),
                 convection_scale=1.0, single_channel=False, conservative=False, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.linear_coefficients = tuple(linear_coefficients)
        self.convection_scale = convection_scale
        self.single_channel = single_channel
        self.conservative = conservative
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt,
                         num_channels=1 if single_channel else num_spatial_dims, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_nonlinear_fun(self, derivative_operator):
        return ConvectionNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=derivative_operator,
                                      dealiasing_fraction=self.dealiasing_fraction, scale=self.convection_scale,
                                      single_channel=self.single_channel, conservative=self.conservative)


class NormalizedConvectionStepper(GeneralConvectionStepper):
    def __init__(self, num_spatial_dims, num_points, *, normalized_linear_coefficients=(0.0, 0.0, 0.01 * 0.1),
                 normalized_convection_scale=1.0 * 0.1, single_channel=False, conservative=False, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        self.normalized_linear_coefficients = tuple(normalized_linear_coefficients)
        self.normalized_convection_scale = normalized_convection_scale
        super().__init__(num_spatial_dims, 1.0, num_points, 1.0, linear_coefficients=normalized_linear_coefficients,
                         convection_scale=normalized_convection_scale, single_channel=single_channel,
                         conservative=conservative, order=order, dealiasing_fraction=dealiasing_fraction,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)


class DifficultyConvectionStepper(NormalizedConvectionStepper):
    def __init__(self, num_spatial_dims=1, num_points=48, *, linear_difficulties=(0.0, 0.0, 4.5),
                 convection_difficulty=5.0, single_channel=False, conservative=False, maximum_absolute=1.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        self.linear_difficulties = tuple(linear_difficulties)
        self.convection_difficulty = convection_difficulty
        kw = dict(num_spatial_dims=num_spatial_dims, num_points=num_points)
        super().__init__(
            num_spatial_dims, num_points,
            normalized_linear_coefficients=u.extract_normalized_coefficients_from_difficulty(linear_difficulties, **kw),
            normalized_convection_scale=u.extract_normalized_convection_scale_from_difficulty(
                convection_difficulty, maximum_absolute=maximum_absolute, **kw),
            single_channel=single_channel, conservative=conservative, order=order,
            dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points, circle_radius=circle_radius)


# ----------------------------------------------------------- gradient norm
class GeneralGradientNormStepper(_Coefficients):
    """``_gradient_norm.py:11-112``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *,
                 linear_coefficients=(0.0, 0.0, -1.0, 0.0, -1.0), gradient_norm_scale=1.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        self.linear_coefficients = tuple(linear_coefficients)
        self.gradient_norm_scale = gradient_norm_scale
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_nonlinear_fun(self, derivative_operator):
        return GradientNormNonlinearFun(self.num_spatial_dims, self.num_points,
                                        derivative_operator=derivative_operator,
                                        dealiasing_fraction=self.dealiasing_fraction, scale=self.gradient_norm_scale,
                                        zero_mode_fix=True)


class NormalizedGradientNormStepper(GeneralGradientNormStepper):
    def __init__(self, num_spatial_dims, num_points, *,
                 normalized_linear_coefficients=(0.0, 0.0, -1.0 * 0.1 / (60.0**2), 0.0, -1.0 * 0.1 / (60.0**4)),
                 normalized_gradient_norm_scale=1.0 * 0.1 / (60.0**2), order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.normalized_linear_coefficients = tuple(normalized_linear_coefficients)
        self.normalized_gradient_norm_scale = normalized_gradient_norm_scale
        super().__init__(num_spatial_dims, 1.0, num_points, 1.0, linear_coefficients=normalized_linear_coefficients,
                         gradient_norm_scale=normalized_gradient_norm_scale, order=order,
                         dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points,
                         circle_radius=circle_radius)


class DifficultyGradientNormStepper(NormalizedGradientNormStepper):
    def __init__(self, num_spatial_dims=1, num_points=48, *, linear_difficulties=(0.0, 0.0, -0.128, 0.0, -0.32768),
                 gradient_norm_difficulty=0.064, maximum_absolute=1.0, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.linear_difficulties = tuple(linear_difficulties)
        self.gradient_norm_difficulty = gradient_norm_difficulty
        kw = dict(num_spatial_dims=num_spatial_dims, num_points=num_points)
        super().__init__(
            num_spatial_dims, num_points,
            normalized_linear_coefficients=u.extract_normalized_coefficients_from_difficulty(linear_difficulties, **kw),
            normalized_gradient_norm_scale=u.extract_normalized_gradient_norm_scale_from_difficulty(
                gradient_norm_difficulty, maximum_absolute=maximum_absolute, **kw),
            order=order, dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points,
            circle_radius=circle_radius)


# -------------------------------------------------------------- polynomial
class GeneralPolynomialStepper(_Coefficients):
    """``_polynomial.py:11-110``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, linear_coefficients=(10.0, 0.0, 1.0),
                 polynomial_coefficients=(0.0, 0.0, -10.0), order=2, dealiasing_fraction=2 / 3, num_circle_points=16,
                 circle_radius=1.0):
        self.linear_coefficients = tuple(linear_coefficients)
        self.polynomial_coefficients = tuple(polynomial_coefficients)
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_nonlinear_fun(self, derivative_operator):
        return PolynomialNonlinearFun(self.num_spatial_dims, self.num_points,
                                      dealiasing_fraction=self.dealiasing_fraction,
                                      coefficients=self.polynomial_coefficients)


class NormalizedPolynomialStepper(GeneralPolynomialStepper):
    def __init__(self, num_spatial_dims, num_points, *,
                 normalized_linear_coefficients=(10.0 * 0.001 / (10.0**0), 0.0, 1.0 * 0.001 / (10.0**2)),
                 normalized_polynomial_coefficients=(0.0, 0.0, -10.0 * 0.001), order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.normalized_linear_coefficients = tuple(normalized_linear_coefficients)
        self.normalized_polynomial_coefficients = tuple(normalized_polynomial_coefficients)
        super().__init__(num_spatial_dims, 1.0, num_points, 1.0, linear_coefficients=normalized_linear_coefficients,
                         polynomial_coefficients=normalized_polynomial_coefficients, order=order,
                         dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points,
                         circle_radius=circle_radius)


class DifficultyPolynomialStepper(NormalizedPolynomialStepper):
    def __init__(self, num_spatial_dims=1, num_points=48, *,
                 linear_difficulties=(10.0 * 0.001 / (10.0**0) * 48**0, 0.0, 1.0 * 0.001 / (10.0**2) * 48**2 * 2**1),
                 polynomial_difficulties=(0.0, 0.0, -10.0 * 0.001), order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.linear_difficulties = tuple(linear_difficulties)
        self.polynomial_difficulties = tuple(polynomial_difficulties)
        super().__init__(
            num_spatial_dims, num_points,
            normalized_linear_coefficients=u.extract_normalized_coefficients_from_difficulty(
                linear_difficulties, num_spatial_dims=num_spatial_dims, num_points=num_points),
            normalized_polynomial_coefficients=polynomial_difficulties,  # difficulties == normalized scales
            order=order, dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points,
            circle_radius=circle_radius)


# --------------------------------------------------------------- nonlinear
class GeneralNonlinearStepper(_Coefficients):
    """``_nonlinear.py:14-120``: b0 u^2 + b1 convection + b2 gradient norm."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, linear_coefficients=(0.0, 0.0, 0.01),
                 nonlinear_coefficients=(0.0, -1.0, 0.0), order=2, dealiasing_fraction=2 / 3, num_circle_points=16,
                 circle_radius=1.0):
        if len(nonlinear_coefficients) != 3:
            raise ValueError("The nonlinear coefficients list must have exactly 3 elements")
        self.linear_coefficients = tuple(linear_coefficients)
        self.nonlinear_coefficients = tuple(nonlinear_coefficients)
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_nonlinear_fun(self, derivative_operator):
        return GeneralNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=derivative_operator,
                                   dealiasing_fraction=self.dealiasing_fraction,
                                   scale_list=self.nonlinear_coefficients, zero_mode_fix=True)


class NormalizedNonlinearStepper(GeneralNonlinearStepper):
    def __init__(self, num_spatial_dims, num_points, *, normalized_linear_coefficients=(0.0, 0.0, 0.1 * 0.1),
                 normalized_nonlinear_coefficients=(0.0, -1.0 * 0.1, 0.0), order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.normalized_linear_coefficients = tuple(normalized_linear_coefficients)
        self.normalized_nonlinear_coefficients = tuple(normalized_nonlinear_coefficients)
        super().__init__(num_spatial_dims, 1.0, num_points, 1.0, linear_coefficients=normalized_linear_coefficients,
                         nonlinear_coefficients=normalized_nonlinear_coefficients, order=order,
                         dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points,
                         circle_radius=circle_radius)


class DifficultyNonlinearStepper(NormalizedNonlinearStepper):
    def __init__(self, num_spatial_dims=1, num_points=48, *,
                 linear_difficulties=(0.0, 0.0, 0.1 * 0.1 / 1.0 * 48**2 * 2),
                 nonlinear_difficulties=(0.0, -1.0 * 0.1 / 1.0 * 48, 0.0), maximum_absolute=1.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        self.linear_difficulties = tuple(linear_difficulties)
        self.nonlinear_difficulties = tuple(nonlinear_difficulties)
        kw = dict(num_spatial_dims=num_spatial_dims, num_points=num_points)
        super().__init__(
            num_spatial_dims, num_points,
            normalized_linear_coefficients=u.extract_normalized_coefficients_from_difficulty(linear_difficulties, **kw),
            normalized_nonlinear_coefficients=u.extract_normalized_nonlinear_scales_from_difficulty(
                nonlinear_difficulties, maximum_absolute=maximum_absolute, **kw),
            order=order, dealiasing_fraction=dealiasing_fraction, num_circle_points=num_circle_points,
            circle_radius=circle_radius)


# ------------------------------------------------------ vorticity (2-D only)
class GeneralVorticityConvectionStepper(_Coefficients):
    """``_vorticity_convection.py:8-142``; with injection_scale != 0 the
    Kolmogorov forcing is added."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, vorticity_convection_scale=1.0,
                 linear_coefficients=(0.0, 0.0, 0.001), injection_mode=4, injection_scale=0.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}.")
        self.vorticity_convection_scale = vorticity_convection_scale
        self.linear_coefficients = tuple(linear_coefficients)
        self.injection_mode = injection_mode
        self.injection_scale = injection_scale
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_nonlinear_fun(self, derivative_operator):
        if self.injection_scale == 0.0:
            return VorticityConvection2d(self.num_spatial_dims, self.num_points,
                                         convection_scale=self.vorticity_convection_scale,
                                         derivative_operator=derivative_operator,
                                         dealiasing_fraction=self.dealiasing_fraction)
        return VorticityConvection2dKolmogorov(self.num_spatial_dims, self.num_points,
                                               convection_scale=self.vorticity_convection_scale,
                                               derivative_operator=derivative_operator,
                                               dealiasing_fraction=self.dealiasing_fraction,
                                               injection_mode=self.injection_mode,
                                               injection_scale=self.injection_scale)
