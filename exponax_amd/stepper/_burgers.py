"""Burgers stepper (``stepper/_burgers.py:8-155``)."""
from .._base_stepper import BaseStepper
from ..nonlin_fun import ConvectionNonlinearFun
from ..spectral import build_laplace_operator


class Burgers(BaseStepper):
    """u_t = nu Laplace u - b (u . grad) u; L = nu sum_d (i kappa_d)^2."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.1, convection_scale=1.0,
                 single_channel=False, conservative=False, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        self.diffusivity = diffusivity
        self.convection_scale = convection_scale
        self.single_channel = single_channel
        self.conservative = conservative
        self.dealiasing_fraction = dealiasing_fraction
        num_channels = 1 if single_channel else num_spatial_dims
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=num_channels, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return self.diffusivity * build_laplace_operator(derivative_operator)

    def _build_nonlinear_fun(self, derivative_operator):
        return ConvectionNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=derivative_operator,
                                      dealiasing_fraction=self.dealiasing_fraction, scale=self.convection_scale,
                                      single_channel=self.single_channel, conservative=self.conservative)
