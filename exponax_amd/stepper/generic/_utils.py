"""Normalized and difficulty parametrisations (``stepper/generic/_utils.py``).

Normalized coefficients fold the domain extent and the time step in:
alpha_j = a_j dt / L^j (a stepper with L = 1, dt = 1); difficulties further
scale by the resolution: gamma_j = alpha_j N^j 2^(j-1) D (gamma_0 = alpha_0),
convection beta = b dt / L -> beta M N D, gradient norm b dt / L^2 ->
beta M N^2 D, with M the maximum absolute value of the state.
"""
from __future__ import annotations


def normalize_coefficients(coefficients, *, domain_extent, dt):
    """``_utils.py:4-37``."""
    return tuple(c * dt / (domain_extent**i) for i, c in enumerate(coefficients))


def denormalize_coefficients(normalized_coefficients, *, domain_extent, dt):
    """``_utils.py:40-68``."""
    return tuple(c / dt * domain_extent**i for i, c in enumerate(normalized_coefficients))


def normalize_convection_scale(convection_scale, *, domain_extent, dt):
    return convection_scale * dt / domain_extent


def denormalize_convection_scale(normalized_convection_scale, *, domain_extent, dt):
    return normalized_convection_scale / dt * domain_extent


def normalize_gradient_norm_scale(gradient_norm_scale, *, domain_extent, dt):
    return gradient_norm_scale * dt / domain_extent**2


def denormalize_gradient_norm_scale(normalized_gradient_norm_scale, *, domain_extent, dt):
    return normalized_gradient_norm_scale / dt * domain_extent**2


def normalize_polynomial_scales(polynomial_scales, *, domain_extent=None, dt):
    return tuple(c * dt for c in polynomial_scales)


def denormalize_polynomial_scales(normalized_polynomial_scales, *, domain_extent=None, dt):
    return tuple(c / dt for c in normalized_polynomial_scales)


def _difficulty_factor(j, num_spatial_dims, num_points):
    return num_points**j * 2 ** (j - 1) * num_spatial_dims


def reduce_normalized_coefficients_to_difficulty(normalized_coefficients, *, num_spatial_dims, num_points):
    """``_utils.py``: gamma_j = alpha_j N^j 2^(j-1) D, gamma_0 = alpha_0."""
    out = [a * _difficulty_factor(j, num_spatial_dims, num_points) for j, a in enumerate(normalized_coefficients)]
    out[0] = normalized_coefficients[0]
    return tuple(out)


def extract_normalized_coefficients_from_difficulty(difficulty_coefficients, *, num_spatial_dims, num_points):
    out = [g / _difficulty_factor(j, num_spatial_dims, num_points) for j, g in enumerate(difficulty_coefficients)]
    out[0] = difficulty_coefficients[0]
    return tuple(out)


def reduce_normalized_convection_scale_to_difficulty(normalized_convection_scale, *, num_spatial_dims, num_points,
                                                     maximum_absolute):
    return normalized_convection_scale * maximum_absolute * num_points * num_spatial_dims


def extract_normalized_convection_scale_from_difficulty(difficulty_convection_scale, *, num_spatial_dims, num_points,
                                                        maximum_absolute):
    return difficulty_convection_scale / (maximum_absolute * num_points * num_spatial_dims)


def reduce_normalized_gradient_norm_scale_to_difficulty(normalized_gradient_norm_scale, *, num_spatial_dims,
                                                        num_points, maximum_absolute):
    return normalized_gradient_norm_scale * maximum_absolute * num_points**2 * num_spatial_dims


def extract_normalized_gradient_norm_scale_from_difficulty(difficulty_gradient_norm_scale, *, num_spatial_dims,
                                                           num_points, maximum_absolute):
    return difficulty_gradient_norm_scale / (maximum_absolute * num_points**2 * num_spatial_dims)


def reduce_normalized_nonlinear_scales_to_difficulty(normalized_nonlinear_scales, *, num_spatial_dims, num_points,
                                                     maximum_absolute):
    kw = dict(num_spatial_dims=num_spatial_dims, num_points=num_points, maximum_absolute=maximum_absolute)
    return (normalized_nonlinear_scales[0],  # polynomial: normalized == difficulty
            reduce_normalized_convection_scale_to_difficulty(normalized_nonlinear_scales[1], **kw),
            reduce_normalized_gradient_norm_scale_to_difficulty(normalized_nonlinear_scales[2], **kw))


def extract_normalized_nonlinear_scales_from_difficulty(nonlinear_difficulties, *, num_spatial_dims, num_points,
                                                        maximum_absolute):
    kw = dict(num_spatial_dims=num_spatial_dims, num_points=num_points, maximum_absolute=maximum_absolute)
    return (nonlinear_difficulties[0],
            extract_normalized_convection_scale_from_difficulty(nonlinear_difficulties[1], **kw),
            extract_normalized_gradient_norm_scale_from_difficulty(nonlinear_difficulties[2], **kw))
