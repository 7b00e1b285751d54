"""Reaction-diffusion steppers with polynomial reactions
(``stepper/reaction/_allen_cahn.py``, ``_fisher_kpp.py``,
``_swift_hohenberg.py``, ``_cahn_hilliard.py``): a real linear operator and
the polynomial nonlinear function (Cahn-Hilliard: the Laplacian of a cubic),
in 1-D and 2-D."""
from .. import _lib
from .._base_stepper import BaseStepper
from ..nonlin_fun import BaseNonlinearFun, PolynomialNonlinearFun
from ..spectral import build_laplace_operator


class _Polynomial(BaseStepper):
    def _build_nonlinear_fun(self, derivative_operator):
        return PolynomialNonlinearFun(self.num_spatial_dims, self.num_points,
                                      dealiasing_fraction=self.dealiasing_fraction,
                                      coefficients=self._poly())


class AllenCahn(_Polynomial):
    """u_t = nu Laplace u + c1 u + c3 u^3 (``_allen_cahn.py``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=5e-3,
                 first_order_coefficient=1.0, third_order_coefficient=-1.0, order=2, dealiasing_fraction=1 / 2,
                 num_circle_points=16, circle_radius=1.0):
        self.diffusivity = diffusivity
        self.first_order_coefficient = first_order_coefficient
        self.third_order_coefficient = third_order_coefficient
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return self.diffusivity * build_laplace_operator(derivative_operator, order=2) + self.first_order_coefficient

    def _poly(self):
        return [0.0, 0.0, 0.0, self.third_order_coefficient]


class FisherKPP(_Polynomial):
    """u_t = nu Laplace u + r u (1 - u) (``_fisher_kpp.py``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01, reactivity=1.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0):
        self.diffusivity = diffusivity
        self.reactivity = reactivity
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        return self.diffusivity * build_laplace_operator(derivative_operator, order=2) + self.reactivity

    def _poly(self):
        return [0.0, 0.0, -self.reactivity]


class SwiftHohenberg(_Polynomial):
    """u_t = r u - (k + Laplace)^2 u + sum_j p_j u^j (``_swift_hohenberg.py``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, reactivity=0.7, critical_number=1.0,
                 polynomial_coefficients=(0.0, 0.0, 1.0, -1.0), order=2, dealiasing_fraction=1 / 2,
                 num_circle_points=16, circle_radius=1.0):
        self.reactivity = reactivity
        self.critical_number = critical_number
        self.polynomial_coefficients = tuple(polynomial_coefficients)
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        lap = build_laplace_operator(derivative_operator, order=2)
        return self.reactivity - (self.critical_number + lap) ** 2

    def _poly(self):
        return list(self.polynomial_coefficients)


class CahnHilliardNonlinearFun(BaseNonlinearFun):
    """``_cahn_hilliard.py:10-41``: scale * Laplace F[u^3] (the state's
    first channel, dealiased before the inverse and after the forward
    transform)."""

    kind = _lib.EXMI_NL_CAHN_HILLIARD

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, scale, dealiasing_fraction):
        from ..nonlin_fun import domain_extent_of

        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.domain_extent = domain_extent_of(derivative_operator, num_spatial_dims, num_points)
        self.scale = scale

    def fill(self, d):
        super().fill(d)
        d.scale = float(self.scale)


class CahnHilliard(BaseStepper):
    """u_t = nu Laplace (c3 u^3 + c1 u - gamma Laplace u) (``_cahn_hilliard.py:44-196``):
    L = nu Laplace (c1 - gamma Laplace), N = nu c3 Laplace F[u^3]."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=1e-2, gamma=1e-3,
                 first_order_coefficient=-1.0, third_order_coefficient=1.0, order=2, dealiasing_fraction=1 / 2,
                 num_circle_points=16, circle_radius=1.0):
        self.diffusivity = diffusivity
        self.gamma = gamma
        self.first_order_coefficient = first_order_coefficient
        self.third_order_coefficient = third_order_coefficient
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        lap = build_laplace_operator(derivative_operator, order=2)
        return self.diffusivity * lap * (self.first_order_coefficient - self.gamma * lap)

    def _build_nonlinear_fun(self, derivative_operator):
        return CahnHilliardNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=derivative_operator,
                                        dealiasing_fraction=self.dealiasing_fraction,
                                        scale=self.diffusivity * self.third_order_coefficient)


class _CoupledReactionFun(BaseNonlinearFun):
    """Reactions that couple the channels (evaluated per grid point on the
    dealiased physical fields, transformed back and dealiased)."""

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction, rates=(0.0, 0.0)):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction)
        self.rates = tuple(float(r) for r in rates)

    def fill(self, d):
        super().fill(d)
        d.reaction_rates[0], d.reaction_rates[1] = self.rates


class GrayScottNonlinearFun(_CoupledReactionFun):
    """``_gray_scott.py:9-49``: (f (1 - u0) - u0 u1^2, -(f + k) u1 + u0 u1^2)."""

    kind = _lib.EXMI_NL_GRAY_SCOTT

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction, feed_rate, kill_rate):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction,
                         rates=(feed_rate, kill_rate))
        self.feed_rate, self.kill_rate = feed_rate, kill_rate


class BelousovZhabotinskyNonlinearFun(_CoupledReactionFun):
    """``_belousov_zhabotinsky.py:9-47``: (u0 + u1 - u0 u1 - u0^2,
    u2 - u1 - u0 u1, u0 - u2)."""

    kind = _lib.EXMI_NL_BELOUSOV_ZHABOTINSKY


class GrayScott(BaseStepper):
    """``_gray_scott.py:52-172``: two channels, L = (nu1 Laplace, nu2 Laplace)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity_1=2e-5, diffusivity_2=1e-5,
                 feed_rate=0.04, kill_rate=0.06, order=2, dealiasing_fraction=1 / 2, num_circle_points=16,
                 circle_radius=1.0):
        self.diffusivity_1, self.diffusivity_2 = diffusivity_1, diffusivity_2
        self.feed_rate, self.kill_rate = feed_rate, kill_rate
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=2, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        import numpy as np

        lap = build_laplace_operator(derivative_operator, order=2)
        return np.concatenate([self.diffusivity_1 * lap, self.diffusivity_2 * lap])

    def _build_nonlinear_fun(self, derivative_operator):
        return GrayScottNonlinearFun(self.num_spatial_dims, self.num_points,
                                     dealiasing_fraction=self.dealiasing_fraction, feed_rate=self.feed_rate,
                                     kill_rate=self.kill_rate)


class BelousovZhabotinsky(BaseStepper):
    """``_belousov_zhabotinsky.py:50-105``: three channels, L = diag(nu_i Laplace)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivities=(1e-5, 2e-5, 1e-5), order=2,
                 dealiasing_fraction=1 / 2, num_circle_points=16, circle_radius=1.0):
        self.diffusivities = tuple(diffusivities)
        self.dealiasing_fraction = dealiasing_fraction
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=3, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius)

    def _build_linear_operator(self, derivative_operator):
        import numpy as np

        lap = build_laplace_operator(derivative_operator, order=2)
        return np.concatenate([nu * lap for nu in self.diffusivities])

    def _build_nonlinear_fun(self, derivative_operator):
        return BelousovZhabotinskyNonlinearFun(self.num_spatial_dims, self.num_points,
                                               dealiasing_fraction=self.dealiasing_fraction)
