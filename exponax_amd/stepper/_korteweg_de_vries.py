"""Korteweg-de Vries stepper (``stepper/_korteweg_de_vries.py:21-224``)."""
import numpy as np

from .._base_stepper import BaseStepper
from ..nonlin_fun import ConvectionNonlinearFun
from ..spectral import build_gradient_inner_product_operator, build_laplace_operator


class KortewegDeVries(BaseStepper):
    """u_t + b/2 (u^2)_x (or b u u_x) + a u_xxx = nu u_xx - nu_h u_xxxx;
    L = nu Laplace - sum_d a (i kappa_d)^3 - nu_h sum_d (i kappa_d)^4 (complex)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, convection_scale=-6.0, diffusivity=0.0,
                 dispersivity=1.0, hyper_diffusivity=0.01, advect_over_diffuse=False, diffuse_over_diffuse=False,
                 single_channel=False, conservative=False, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.convection_scale = convection_scale
        self.diffusivity = diffusivity
        self.dispersivity = dispersivity
        self.hyper_diffusivity = hyper_diffusivity
        self.advect_over_diffuse = advect_over_diffuse
        self.diffuse_over_diffuse = diffuse_over_diffuse
        self.single_channel = single_channel
        self.conservative = conservative
        self.dealiasing_fraction = dealiasing_fraction
        num_channels = 1 if single_channel else num_spatial_dims
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=num_channels, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        # _korteweg_de_vries.py:170-205
        velocity = self.dispersivity * np.ones(self.num_spatial_dims)
        lap = build_laplace_operator(derivative_operator, order=2)
        if self.advect_over_diffuse:
            dispersion = -build_gradient_inner_product_operator(derivative_operator, velocity, order=1) * lap
        else:
            dispersion = -build_gradient_inner_product_operator(derivative_operator, velocity, order=3)
        if self.diffuse_over_diffuse:
            hyper = -self.hyper_diffusivity * lap * lap
        else:
            hyper = -self.hyper_diffusivity * build_laplace_operator(derivative_operator, order=4)
        return self.diffusivity * lap + dispersion + hyper

    def _build_nonlinear_fun(self, derivative_operator):
        return ConvectionNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=derivative_operator,
                                      dealiasing_fraction=self.dealiasing_fraction, scale=self.convection_scale,
                                      single_channel=self.single_channel, conservative=self.conservative)
