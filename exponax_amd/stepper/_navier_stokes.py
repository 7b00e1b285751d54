"""2-D vorticity steppers (``stepper/_navier_stokes.py:13-327``)."""
from .._base_stepper import BaseStepper
from ..nonlin_fun import (ProjectedConvection3d, ProjectedConvection3dKolmogorov, VorticityConvection2d,
                          VorticityConvection2dKolmogorov)
from ..spectral import build_laplace_operator


class NavierStokesVorticity(BaseStepper):
    """``_navier_stokes.py:13-153``: L = nu Laplace + drag."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01,
                 vorticity_convection_scale=1.0, drag=0.0, order=2, dealiasing_fraction=2 / 3, num_circle_points=16,
                 circle_radius=1.0):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}. For 3D, use "
                             "NavierStokesVelocity instead.")
        self.diffusivity = diffusivity
        self.vorticity_convection_scale = vorticity_convection_scale
        self.drag = drag
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return (self.diffusivity * build_laplace_operator(derivative_operator, order=2)
                + self.drag * build_laplace_operator(derivative_operator, order=0))

    def _build_nonlinear_fun(self, derivative_operator):
        return VorticityConvection2d(self.num_spatial_dims, self.num_points,
                                     convection_scale=self.vorticity_convection_scale,
                                     derivative_operator=derivative_operator,
                                     dealiasing_fraction=self.dealiasing_fraction)


class KolmogorovFlowVorticity(BaseStepper):
    """``_navier_stokes.py:156-327``: adds the forcing -k_f gamma sin(k_f y)
    in vorticity form."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.001, convection_scale=1.0,
                 drag=-0.1, injection_mode=4, injection_scale=1.0, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}. For 3D, use "
                             "KolmogorovFlowVelocity instead.")
        self.diffusivity = diffusivity
        self.convection_scale = convection_scale
        self.drag = drag
        self.injection_mode = injection_mode
        self.injection_scale = injection_scale
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return (self.diffusivity * build_laplace_operator(derivative_operator, order=2)
                + self.drag * build_laplace_operator(derivative_operator, order=0))

    def _build_nonlinear_fun(self, derivative_operator):
        return VorticityConvection2dKolmogorov(self.num_spatial_dims, self.num_points,
                                               convection_scale=self.convection_scale,
                                               injection_mode=self.injection_mode,
                                               injection_scale=self.injection_scale,
                                               derivative_operator=derivative_operator,
                                               dealiasing_fraction=self.dealiasing_fraction)


class NavierStokesVelocity(BaseStepper):
    """``stepper/_navier_stokes.py:330-450``: 3-D incompressible Navier-Stokes
    in velocity form, u_t = P(u x omega) + nu Laplace u + lambda u (three
    channels)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01, drag=0.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        if num_spatial_dims != 3:
            raise ValueError(f"NavierStokesVelocity is three-dimensional, got num_spatial_dims={num_spatial_dims}")
        self.diffusivity = diffusivity
        self.drag = drag
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=3, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return (self.diffusivity * build_laplace_operator(derivative_operator, order=2)
                + self.drag * build_laplace_operator(derivative_operator, order=0))

    def _build_nonlinear_fun(self, derivative_operator):
        return ProjectedConvection3d(self.num_spatial_dims, self.num_points, derivative_operator=derivative_operator,
                                     dealiasing_fraction=self.dealiasing_fraction)


class KolmogorovFlowVelocity(BaseStepper):
    """``stepper/_navier_stokes.py:451-598``: NavierStokesVelocity with the
    Kolmogorov forcing in the first channel."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01, drag=0.0,
                 injection_mode=4, injection_scale=1.0, order=2, dealiasing_fraction=2 / 3, num_circle_points=16,
                 circle_radius=1.0):
        if num_spatial_dims != 3:
            raise ValueError(f"KolmogorovFlowVelocity is three-dimensional, got num_spatial_dims={num_spatial_dims}")
        self.diffusivity = diffusivity
        self.drag = drag
        self.injection_mode = injection_mode
        self.injection_scale = injection_scale
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=3, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return (self.diffusivity * build_laplace_operator(derivative_operator, order=2)
                + self.drag * build_laplace_operator(derivative_operator, order=0))

    def _build_nonlinear_fun(self, derivative_operator):
        return ProjectedConvection3dKolmogorov(self.num_spatial_dims, self.num_points,
                                               injection_mode=self.injection_mode,
                                               injection_scale=self.injection_scale,
                                               derivative_operator=derivative_operator,
                                               dealiasing_fraction=self.dealiasing_fraction)
