"""Linear steppers: advection, diffusion, advection-diffusion, dispersion and
hyper-diffusion (``stepper/_advection.py``, ``_diffusion.py``,
``_advection_diffusion.py``, ``_dispersion.py``, ``_hyper_diffusion.py``).

All are ETDRK0 (the exact exponential) with a zero nonlinear term. The odd
derivatives make L complex; those run the complex-operator kernels (1-D).
"""
import numpy as np

from .._base_stepper import BaseStepper
from ..nonlin_fun import ZeroNonlinearFun
from ..spectral import build_gradient_inner_product_operator, build_laplace_operator


def _vector(value, num_spatial_dims):
    # _advection.py:80-82: a float is broadcast to every axis
    return np.ones(num_spatial_dims) * value if np.isscalar(value) else np.asarray(value, dtype=np.float64)


def _diffusivity_matrix(value, num_spatial_dims):
    # _diffusion.py:84-88: float -> nu I, vector -> diag, matrix kept
    if np.isscalar(value):
        return np.eye(num_spatial_dims) * value
    value = np.asarray(value, dtype=np.float64)
    return np.diag(value) if value.ndim == 1 else value


def _anisotropic_laplace(derivative_operator, matrix):
    # sum_ij A_ij (i kappa_i)(i kappa_j)  (_diffusion.py:99-113)
    outer = derivative_operator[:, None] * derivative_operator[None, :]
    return np.einsum("ij,ij...->...", matrix, outer)[None]


class _Linear(BaseStepper):
    def __init__(self, num_spatial_dims, domain_extent, num_points, dt):
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=0)

    def _build_nonlinear_fun(self, derivative_operator):
        return ZeroNonlinearFun(self.num_spatial_dims, self.num_points)


class Advection(_Linear):
    """u_t + c . grad u = 0; L = -sum_d c_d (i kappa_d) (``_advection.py:15-104``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, velocity=1.0):
        self.velocity = _vector(velocity, num_spatial_dims)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt)

    def _build_linear_operator(self, derivative_operator):
        return -build_gradient_inner_product_operator(derivative_operator, self.velocity, order=1)


class Diffusion(_Linear):
    """u_t = div(A grad u); L = sum_ij A_ij (i kappa_i)(i kappa_j) (``_diffusion.py:13-122``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01):
        self.diffusivity = _diffusivity_matrix(diffusivity, num_spatial_dims)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt)

    def _build_linear_operator(self, derivative_operator):
        return _anisotropic_laplace(derivative_operator, self.diffusivity)


class AdvectionDiffusion(_Linear):
    """u_t + c . grad u = div(A grad u) (``_advection_diffusion.py:15-136``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, velocity=1.0, diffusivity=0.01):
        self.velocity = _vector(velocity, num_spatial_dims)
        self.diffusivity = _diffusivity_matrix(diffusivity, num_spatial_dims)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt)

    def _build_linear_operator(self, derivative_operator):
        return (_anisotropic_laplace(derivative_operator, self.diffusivity)
                - build_gradient_inner_product_operator(derivative_operator, self.velocity, order=1))


class Dispersion(_Linear):
    """u_t = sum_d a_d d^3 u / dx_d^3 (or a . grad Laplace u with
    advect_on_diffusion) (``_dispersion.py:15-126``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, dispersivity=1.0,
                 advect_on_diffusion=False):
        self.dispersivity = _vector(dispersivity, num_spatial_dims)
        self.advect_on_diffusion = advect_on_diffusion
        super().__init__(num_spatial_dims, domain_extent, num_points, dt)

    def _build_linear_operator(self, derivative_operator):
        if self.advect_on_diffusion:
            return (build_gradient_inner_product_operator(derivative_operator, self.dispersivity, order=1)
                    * build_laplace_operator(derivative_operator))
        return build_gradient_inner_product_operator(derivative_operator, self.dispersivity, order=3)


class HyperDiffusion(_Linear):
    """u_t = -nu_h sum_d d^4 u / dx_d^4 (or -nu_h Laplace^2 u)
    (``_hyper_diffusion.py:10-120``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, hyper_diffusivity=0.0001,
                 diffuse_on_diffuse=False):
        self.hyper_diffusivity = hyper_diffusivity
        self.diffuse_on_diffuse = diffuse_on_diffuse
        super().__init__(num_spatial_dims, domain_extent, num_points, dt)

    def _build_linear_operator(self, derivative_operator):
        if self.diffuse_on_diffuse:
            lap = build_laplace_operator(derivative_operator)
            return -self.hyper_diffusivity * lap * lap
        return -self.hyper_diffusivity * build_laplace_operator(derivative_operator, order=4)
