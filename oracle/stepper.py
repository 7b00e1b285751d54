"""CPU restatement of exponax's stepper layer (test infrastructure only).

``BaseStepper`` follows ``exponax/_base_stepper.py:17-271``; the concrete
steppers follow ``exponax/stepper/_kuramoto_sivashinsky.py``,
``_burgers.py``, ``_navier_stokes.py`` (vorticity and velocity forms), the
linear steppers (``_advection.py`` ... ``_korteweg_de_vries.py``) and
``stepper/reaction/``; ``rollout``/``repeat``
follow ``exponax/_utils.py:92-254``; ``nrmse`` follows
``exponax/metrics/_spatial.py:537-586``.

Batching: a stepper maps a single state (C, N, ..., N). ``step_batch``
applies it to (B, C, N, ..., N) by running the same vectorised NumPy code
with a leading batch axis — identical arithmetic to a vmap.
"""
from __future__ import annotations

import numpy as np

from . import etdrk as et
from . import nonlin as nl
from . import spectral as sp


class BaseStepper:
    """``_base_stepper.py:17-157``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, num_channels, order,
                 num_circle_points=16, circle_radius=1.0, dtype=np.float64, coef_dtype=np.float64):
        self.num_spatial_dims = num_spatial_dims
        self.domain_extent = domain_extent
        self.num_points = num_points
        self.dt = dt
        self.num_channels = num_channels
        self.dx = domain_extent / num_points
        self.dtype = np.dtype(dtype)
        deriv = sp.build_derivative_operator(num_spatial_dims, domain_extent, num_points)
        lin = self._build_linear_operator(deriv)
        single = (1,) + sp.wavenumber_shape(num_spatial_dims, num_points)
        multi = (num_channels,) + sp.wavenumber_shape(num_spatial_dims, num_points)
        if lin.shape not in (single, multi):
            raise ValueError(f"Expected linear operator to have shape {single} or {multi}, got {lin.shape}.")
        self.linear_operator = lin
        nonlin = self._build_nonlinear_fun(deriv)
        kw = dict(dtype=dtype, coef_dtype=coef_dtype)
        ckw = dict(num_circle_points=num_circle_points, circle_radius=circle_radius)
        if order == 0:
            self._integrator = et.ETDRK0(dt, lin, **kw)
        elif order == 1:
            self._integrator = et.ETDRK1(dt, lin, nonlin, **ckw, **kw)
        elif order == 2:
            self._integrator = et.ETDRK2(dt, lin, nonlin, **ckw, **kw)
        elif order == 3:
            self._integrator = et.ETDRK3(dt, lin, nonlin, **ckw, **kw)
        elif order == 4:
            self._integrator = et.ETDRK4(dt, lin, nonlin, **ckw, **kw)
        else:
            raise NotImplementedError(f"Order {order} not implemented.")
        self.nonlinear_fun = nonlin

    def step(self, u):
        """``_base_stepper.py:201-220``: fft -> step_fourier -> ifft."""
        u = np.asarray(u, dtype=self.dtype)
        u_hat = sp.fft(u, num_spatial_dims=self.num_spatial_dims).astype(sp.complex_of(self.dtype))
        nxt = self.step_fourier(u_hat)
        out = sp.ifft(nxt, num_spatial_dims=self.num_spatial_dims, num_points=self.num_points)
        return out.astype(self.dtype, copy=False)

    def step_fourier(self, u_hat):
        return self._integrator.step_fourier(u_hat)

    def __call__(self, u):
        """``_base_stepper.py:241-271``."""
        expected = (self.num_channels,) + sp.spatial_shape(self.num_spatial_dims, self.num_points)
        if u.shape != expected:
            raise ValueError(f"Expected shape {expected}, got {u.shape}. For batched operation use "
                             "`jax.vmap` on this function.")
        return self.step(u)

    def step_batch(self, ub):
        """Vectorised step over a leading batch axis (the vmap of ``__call__``).

        Every operator broadcasts over a leading axis because the reference's
        operators carry a leading channel axis of size 1 or C; a batch axis in
        front of that broadcasts identically.
        """
        return self.step(ub)


class KuramotoSivashinskyConservative(BaseStepper):
    """``stepper/_kuramoto_sivashinsky.py:171-315``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, convection_scale=1.0,
                 second_order_scale=1.0, fourth_order_scale=1.0, single_channel=False, conservative=True,
                 dealiasing_fraction=2 / 3, order=2, num_circle_points=16, circle_radius=1.0, **kw):
        self.convection_scale = convection_scale
        self.second_order_scale = second_order_scale
        self.fourth_order_scale = fourth_order_scale
        self.single_channel = single_channel
        self.conservative = conservative
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        nc = 1 if single_channel else num_spatial_dims
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=nc, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return (-self.second_order_scale * sp.build_laplace_operator(d, order=2)
                - self.fourth_order_scale * sp.build_laplace_operator(d, order=4))

    def _build_nonlinear_fun(self, d):
        return nl.ConvectionNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                         dealiasing_fraction=self.dealiasing_fraction,
                                         scale=self.convection_scale, single_channel=self.single_channel,
                                         conservative=self.conservative, dtype=self._kw_dtype)


class KuramotoSivashinsky(BaseStepper):
    """``stepper/_kuramoto_sivashinsky.py:8-168`` (combustion / gradient-norm form)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, gradient_norm_scale=1.0,
                 second_order_scale=1.0, fourth_order_scale=1.0, dealiasing_fraction=2 / 3, order=2,
                 num_circle_points=16, circle_radius=1.0, **kw):
        self.gradient_norm_scale = gradient_norm_scale
        self.second_order_scale = second_order_scale
        self.fourth_order_scale = fourth_order_scale
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return (-self.second_order_scale * sp.build_laplace_operator(d, order=2)
                - self.fourth_order_scale * sp.build_laplace_operator(d, order=4))

    def _build_nonlinear_fun(self, d):
        return nl.GradientNormNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                           dealiasing_fraction=self.dealiasing_fraction, zero_mode_fix=True,
                                           scale=self.gradient_norm_scale, dtype=self._kw_dtype)


class Burgers(BaseStepper):
    """``stepper/_burgers.py:8-155``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.1,
                 convection_scale=1.0, single_channel=False, conservative=False, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0, **kw):
        self.diffusivity = diffusivity
        self.convection_scale = convection_scale
        self.single_channel = single_channel
        self.conservative = conservative
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        nc = 1 if single_channel else num_spatial_dims
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=nc, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return self.diffusivity * sp.build_laplace_operator(d)

    def _build_nonlinear_fun(self, d):
        return nl.ConvectionNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                         dealiasing_fraction=self.dealiasing_fraction,
                                         scale=self.convection_scale, single_channel=self.single_channel,
                                         conservative=self.conservative, dtype=self._kw_dtype)


def _vec(value, d):
    return np.ones(d) * value if np.isscalar(value) else np.asarray(value, dtype=np.float64)


def _diff_matrix(value, d):
    if np.isscalar(value):
        return np.eye(d) * value
    value = np.asarray(value, dtype=np.float64)
    return np.diag(value) if value.ndim == 1 else value


class _LinearStepper(BaseStepper):
    """ETDRK0 with a zero nonlinear term (the linear steppers of
    ``stepper/_advection.py``, ``_diffusion.py``, ``_advection_diffusion.py``,
    ``_dispersion.py``, ``_hyper_diffusion.py``)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, **kw):
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=0, **kw)

    def _build_nonlinear_fun(self, d):
        return nl.ZeroNonlinearFun(self.num_spatial_dims, self.num_points)


class Advection(_LinearStepper):
    """``stepper/_advection.py:15-104``: L = -sum_d c_d (i kappa_d)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, velocity=1.0, **kw):
        self.velocity = _vec(velocity, num_spatial_dims)  # :80-82
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, **kw)

    def _build_linear_operator(self, d):
        return -sp.build_gradient_inner_product_operator(d, self.velocity, order=1)  # :88-95


class Diffusion(_LinearStepper):
    """``stepper/_diffusion.py:13-122``: L = sum_ij A_ij (i kappa_i)(i kappa_j)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01, **kw):
        self.diffusivity = _diff_matrix(diffusivity, num_spatial_dims)  # :84-88
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, **kw)

    def _build_linear_operator(self, d):
        return np.einsum("ij,ij...->...", self.diffusivity, d[:, None] * d[None, :])[None]  # :99-113


class AdvectionDiffusion(_LinearStepper):
    """``stepper/_advection_diffusion.py:15-136``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, velocity=1.0, diffusivity=0.01, **kw):
        self.velocity = _vec(velocity, num_spatial_dims)
        self.diffusivity = _diff_matrix(diffusivity, num_spatial_dims)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, **kw)

    def _build_linear_operator(self, d):  # :107-128
        diffusion = np.einsum("ij,ij...->...", self.diffusivity, d[:, None] * d[None, :])[None]
        return -sp.build_gradient_inner_product_operator(d, self.velocity, order=1) + diffusion


class Dispersion(_LinearStepper):
    """``stepper/_dispersion.py:15-126``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, dispersivity=1.0,
                 advect_on_diffusion=False, **kw):
        self.dispersivity = _vec(dispersivity, num_spatial_dims)
        self.advect_on_diffusion = advect_on_diffusion
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, **kw)

    def _build_linear_operator(self, d):  # :102-117
        if self.advect_on_diffusion:
            return (sp.build_gradient_inner_product_operator(d, self.dispersivity, order=1)
                    * sp.build_laplace_operator(d))
        return sp.build_gradient_inner_product_operator(d, self.dispersivity, order=3)


class HyperDiffusion(_LinearStepper):
    """``stepper/_hyper_diffusion.py:10-120``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, hyper_diffusivity=0.0001,
                 diffuse_on_diffuse=False, **kw):
        self.hyper_diffusivity = hyper_diffusivity
        self.diffuse_on_diffuse = diffuse_on_diffuse
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, **kw)

    def _build_linear_operator(self, d):  # :95-110
        if self.diffuse_on_diffuse:
            lap = sp.build_laplace_operator(d)
            return -self.hyper_diffusivity * lap * lap
        return -self.hyper_diffusivity * sp.build_laplace_operator(d, order=4)


class KortewegDeVries(BaseStepper):
    """``stepper/_korteweg_de_vries.py:21-224``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, convection_scale=-6.0, diffusivity=0.0,
                 dispersivity=1.0, hyper_diffusivity=0.01, advect_over_diffuse=False, diffuse_over_diffuse=False,
                 single_channel=False, conservative=False, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0, **kw):
        self.convection_scale = convection_scale
        self.diffusivity = diffusivity
        self.dispersivity = dispersivity
        self.hyper_diffusivity = hyper_diffusivity
        self.advect_over_diffuse = advect_over_diffuse
        self.diffuse_over_diffuse = diffuse_over_diffuse
        self.single_channel = single_channel
        self.conservative = conservative
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        nc = 1 if single_channel else num_spatial_dims
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=nc, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):  # :170-205
        velocity = self.dispersivity * np.ones(self.num_spatial_dims)
        lap = sp.build_laplace_operator(d, order=2)
        if self.advect_over_diffuse:
            dispersion = -sp.build_gradient_inner_product_operator(d, velocity, order=1) * lap
        else:
            dispersion = -sp.build_gradient_inner_product_operator(d, velocity, order=3)
        if self.diffuse_over_diffuse:
            hyper = -self.hyper_diffusivity * lap * lap
        else:
            hyper = -self.hyper_diffusivity * sp.build_laplace_operator(d, order=4)
        return self.diffusivity * lap + dispersion + hyper

    def _build_nonlinear_fun(self, d):  # :207-224
        return nl.ConvectionNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                         dealiasing_fraction=self.dealiasing_fraction,
                                         scale=self.convection_scale, single_channel=self.single_channel,
                                         conservative=self.conservative, dtype=self._kw_dtype)


class _PolynomialReaction(BaseStepper):
    """Real linear operator + ``PolynomialNonlinearFun`` (``stepper/reaction/``)."""

    def _build_nonlinear_fun(self, d):
        return nl.PolynomialNonlinearFun(self.num_spatial_dims, self.num_points,
                                         dealiasing_fraction=self.dealiasing_fraction, coefficients=self._poly(),
                                         dtype=self._kw_dtype)


class AllenCahn(_PolynomialReaction):
    """``stepper/reaction/_allen_cahn.py``: L = nu Laplace + c1, N = c3 u^3."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=5e-3,
                 first_order_coefficient=1.0, third_order_coefficient=-1.0, order=2, dealiasing_fraction=1 / 2,
                 num_circle_points=16, circle_radius=1.0, **kw):
        self.diffusivity, self.c1, self.c3 = diffusivity, first_order_coefficient, third_order_coefficient
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return self.diffusivity * sp.build_laplace_operator(d, order=2) + self.c1

    def _poly(self):
        return [0.0, 0.0, 0.0, self.c3]


class FisherKPP(_PolynomialReaction):
    """``stepper/reaction/_fisher_kpp.py``: L = nu Laplace + r, N = -r u^2."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01, reactivity=1.0,
                 order=2, dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0, **kw):
        self.diffusivity, self.reactivity = diffusivity, reactivity
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return self.diffusivity * sp.build_laplace_operator(d, order=2) + self.reactivity

    def _poly(self):
        return [0.0, 0.0, -self.reactivity]


class SwiftHohenberg(_PolynomialReaction):
    """``stepper/reaction/_swift_hohenberg.py``: L = r - (k + Laplace)^2."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, reactivity=0.7, critical_number=1.0,
                 polynomial_coefficients=(0.0, 0.0, 1.0, -1.0), order=2, dealiasing_fraction=1 / 2,
                 num_circle_points=16, circle_radius=1.0, **kw):
        self.reactivity, self.critical_number = reactivity, critical_number
        self.polynomial_coefficients = tuple(polynomial_coefficients)
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return self.reactivity - (self.critical_number + sp.build_laplace_operator(d, order=2)) ** 2

    def _poly(self):
        return list(self.polynomial_coefficients)


class CahnHilliard(BaseStepper):
    """``stepper/reaction/_cahn_hilliard.py:44-196``: L = nu Laplace (c1 -
    gamma Laplace), N = nu c3 Laplace F[u^3]."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=1e-2, gamma=1e-3,
                 first_order_coefficient=-1.0, third_order_coefficient=1.0, order=2, dealiasing_fraction=1 / 2,
                 num_circle_points=16, circle_radius=1.0, **kw):
        self.diffusivity, self.gamma = diffusivity, gamma
        self.c1, self.c3 = first_order_coefficient, third_order_coefficient
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        lap = sp.build_laplace_operator(d, order=2)
        return self.diffusivity * lap * (self.c1 - self.gamma * lap)

    def _build_nonlinear_fun(self, d):
        return nl.CahnHilliardNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                           scale=self.diffusivity * self.c3,
                                           dealiasing_fraction=self.dealiasing_fraction, dtype=self._kw_dtype)


class GrayScott(BaseStepper):
    """``stepper/reaction/_gray_scott.py:52-172``: L = (nu1 Laplace, nu2 Laplace)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity_1=2e-5, diffusivity_2=1e-5,
                 feed_rate=0.04, kill_rate=0.06, order=2, dealiasing_fraction=1 / 2, num_circle_points=16,
                 circle_radius=1.0, **kw):
        self.nu1, self.nu2, self.feed_rate, self.kill_rate = diffusivity_1, diffusivity_2, feed_rate, kill_rate
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=2, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        lap = sp.build_laplace_operator(d, order=2)
        return np.concatenate([self.nu1 * lap, self.nu2 * lap])

    def _build_nonlinear_fun(self, d):
        return nl.GrayScottNonlinearFun(self.num_spatial_dims, self.num_points,
                                        dealiasing_fraction=self.dealiasing_fraction, feed_rate=self.feed_rate,
                                        kill_rate=self.kill_rate, dtype=self._kw_dtype)


class BelousovZhabotinsky(BaseStepper):
    """``stepper/reaction/_belousov_zhabotinsky.py:50-105``: L = diag(nu_i Laplace)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivities=(1e-5, 2e-5, 1e-5), order=2,
                 dealiasing_fraction=1 / 2, num_circle_points=16, circle_radius=1.0, **kw):
        self.diffusivities = tuple(diffusivities)
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=3, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        lap = sp.build_laplace_operator(d, order=2)
        return np.concatenate([nu * lap for nu in self.diffusivities])

    def _build_nonlinear_fun(self, d):
        return nl.BelousovZhabotinskyNonlinearFun(self.num_spatial_dims, self.num_points,
                                                  dealiasing_fraction=self.dealiasing_fraction,
                                                  dtype=self._kw_dtype)


class NavierStokesVorticity(BaseStepper):
    """``stepper/_navier_stokes.py:13-153``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01,
                 vorticity_convection_scale=1.0, drag=0.0, order=2, dealiasing_fraction=2 / 3,
                 num_circle_points=16, circle_radius=1.0, **kw):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}.")
        self.diffusivity = diffusivity
        self.vorticity_convection_scale = vorticity_convection_scale
        self.drag = drag
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return (self.diffusivity * sp.build_laplace_operator(d, order=2)
                + self.drag * sp.build_laplace_operator(d, order=0))

    def _build_nonlinear_fun(self, d):
        return nl.VorticityConvection2d(self.num_spatial_dims, self.num_points,
                                        convection_scale=self.vorticity_convection_scale, derivative_operator=d,
                                        dealiasing_fraction=self.dealiasing_fraction, dtype=self._kw_dtype)


class KolmogorovFlowVorticity(BaseStepper):
    """``stepper/_navier_stokes.py:156-327``."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.001,
                 convection_scale=1.0, drag=-0.1, injection_mode=4, injection_scale=1.0, order=2,
                 dealiasing_fraction=2 / 3, num_circle_points=16, circle_radius=1.0, **kw):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}.")
        self.diffusivity = diffusivity
        self.convection_scale = convection_scale
        self.drag = drag
        self.injection_mode = injection_mode
        self.injection_scale = injection_scale
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return (self.diffusivity * sp.build_laplace_operator(d, order=2)
                + self.drag * sp.build_laplace_operator(d, order=0))

    def _build_nonlinear_fun(self, d):
        return nl.VorticityConvection2dKolmogorov(
            self.num_spatial_dims, self.num_points, convection_scale=self.convection_scale,
            injection_mode=self.injection_mode, injection_scale=self.injection_scale, derivative_operator=d,
            dealiasing_fraction=self.dealiasing_fraction, dtype=self._kw_dtype)


class NavierStokesVelocity(BaseStepper):
    """``stepper/_navier_stokes.py:330-450``: L = nu Laplace + lambda,
    N = P(u x omega) (+ the Kolmogorov forcing with injection_mode set)."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, diffusivity=0.01, drag=0.0,
                 injection_mode=None, injection_scale=1.0, order=2, dealiasing_fraction=2 / 3, num_circle_points=16,
                 circle_radius=1.0, **kw):
        self.diffusivity, self.drag = diffusivity, drag
        self.injection_mode, self.injection_scale = injection_mode, injection_scale
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=3, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        return self.diffusivity * sp.build_laplace_operator(d, order=2) + self.drag * sp.build_laplace_operator(
            d, order=0)

    def _build_nonlinear_fun(self, d):
        if self.injection_mode is None:
            return nl.ProjectedConvection3d(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                            dealiasing_fraction=self.dealiasing_fraction, dtype=self._kw_dtype)
        return nl.ProjectedConvection3dKolmogorov(self.num_spatial_dims, self.num_points,
                                                  injection_mode=self.injection_mode,
                                                  injection_scale=self.injection_scale, derivative_operator=d,
                                                  dealiasing_fraction=self.dealiasing_fraction, dtype=self._kw_dtype)


class Wave:
    """``stepper/_wave.py:9-197``: h_t = v, v_t = c^2 Laplace h, stepped by
    the reference's diagonalisation: w = i c |k| h (|k| -> 1 at the mean
    mode), modes (w +- v)/sqrt 2 advanced by exp(+- i c |k| dt) (ETDRK0 with
    L = (i c |k|, -i c |k|)), rotated back, h = w / (i c |k|), and the mean
    mode's drift h += dt v."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, speed_of_sound=1.0, dtype=np.float64):
        self.num_spatial_dims, self.num_points, self.dt = num_spatial_dims, num_points, dt
        self.speed_of_sound = speed_of_sound
        self.dtype = np.dtype(dtype)
        cdt = sp.complex_of(dtype)
        k = sp.build_wavenumbers(num_spatial_dims, num_points) * (2 * np.pi / domain_extent)
        self.knorm = np.linalg.norm(k, axis=0, keepdims=True)
        lin = 1j * speed_of_sound * self.knorm
        self.exp_term = np.exp(dt * np.concatenate([lin, -lin])).astype(cdt)
        self.kguard = np.where(self.knorm == 0, 1.0, self.knorm)

    def step_fourier(self, u_hat):
        cdt = u_hat.dtype
        ax = u_hat.ndim - self.num_spatial_dims - 1
        h, v = np.take(u_hat, [0], axis=ax), np.take(u_hat, [1], axis=ax)
        w = (1j * self.speed_of_sound * self.kguard).astype(cdt) * h
        r = cdt.type(1 / np.sqrt(2))
        waves = np.concatenate([r * (w + v), r * (w - v)], axis=ax) * self.exp_term
        p, q = np.take(waves, [0], axis=ax), np.take(waves, [1], axis=ax)
        w1, v1 = r * (p + q), r * (p - q)
        h1 = w1 / (1j * self.speed_of_sound * self.kguard).astype(cdt)
        out = np.concatenate([h1, v1], axis=ax)
        dc = (slice(None),) * ax + (0,) + (0,) * self.num_spatial_dims
        dcv = (slice(None),) * ax + (1,) + (0,) * self.num_spatial_dims
        out[dc] = out[dc] + cdt.type(self.dt) * u_hat[dcv]
        return out

    def step_batch(self, u):
        u = np.asarray(u, dtype=self.dtype)
        uh = sp.fft(u, num_spatial_dims=self.num_spatial_dims).astype(sp.complex_of(self.dtype))
        return sp.ifft(self.step_fourier(uh), num_spatial_dims=self.num_spatial_dims,
                       num_points=self.num_points).astype(self.dtype, copy=False)

    step = step_batch


class Poisson:
    """``_poisson.py:12-80``: u_hat = -f_hat / Laplace_order, 0 where the
    operator vanishes."""

    def __init__(self, num_spatial_dims, domain_extent, num_points, *, order=2, dtype=np.float64):
        self.num_spatial_dims, self.num_points, self.dtype = num_spatial_dims, num_points, np.dtype(dtype)
        d = sp.build_derivative_operator(num_spatial_dims, domain_extent, num_points)
        op = sp.build_laplace_operator(d, order=order)
        with np.errstate(divide="ignore", invalid="ignore"):
            self.inv = np.where(op == 0, 0.0, 1.0 / np.where(op == 0, 1.0, op)).astype(sp.complex_of(dtype))

    def step_fourier(self, f_hat):
        return -self.inv * f_hat

    def step(self, f):
        f = np.asarray(f, dtype=self.dtype)
        fh = sp.fft(f, num_spatial_dims=self.num_spatial_dims).astype(sp.complex_of(self.dtype))
        return sp.ifft(self.step_fourier(fh), num_spatial_dims=self.num_spatial_dims,
                       num_points=self.num_points).astype(self.dtype, copy=False)


class ForcedStepper:
    """``_forced_stepper.py:8-56``: stepper(u + dt f)."""

    def __init__(self, stepper):
        self.stepper = stepper

    def step_batch(self, u, f):
        return self.stepper.step_batch(u + self.stepper.dt * f)


class GeneralStepper(BaseStepper):
    """Linear part sum_j a_j sum_d (i kappa_d)^j (even j here) + ``GeneralNonlinearFun``.

    Mirrors the stepper families of ``stepper/generic/_convection.py`` /
    ``_nonlinear.py`` restricted to even-order (real) linear terms.
    """

    def __init__(self, num_spatial_dims, domain_extent, num_points, dt, *, linear_coefficients=(0.0, 0.0, 0.01),
                 nonlinear_scales=(0.0, -1.0, 0.0), dealiasing_fraction=2 / 3, order=2, num_circle_points=16,
                 circle_radius=1.0, **kw):
        self.linear_coefficients = tuple(linear_coefficients)
        self.nonlinear_scales = tuple(nonlinear_scales)
        self.dealiasing_fraction = dealiasing_fraction
        self._kw_dtype = kw.get("dtype", np.float64)
        super().__init__(num_spatial_dims, domain_extent, num_points, dt, num_channels=1, order=order,
                         num_circle_points=num_circle_points, circle_radius=circle_radius, **kw)

    def _build_linear_operator(self, d):
        out = np.zeros((1,) + d.shape[1:], dtype=np.complex128)
        for j, a in enumerate(self.linear_coefficients):
            out = out + a * sp.build_laplace_operator(d, order=j) if j % 2 == 0 else out + a * np.sum(
                d**j, axis=0, keepdims=True)
        return out

    def _build_nonlinear_fun(self, d):
        return nl.GeneralNonlinearFun(self.num_spatial_dims, self.num_points, derivative_operator=d,
                                      dealiasing_fraction=self.dealiasing_fraction,
                                      scale_list=self.nonlinear_scales, dtype=self._kw_dtype)


def rollout(stepper_fn, n, *, include_init=False):
    """``_utils.py:92-186`` (takes_aux=False): stack of n successive states."""
    def run(u0):
        out = [u0] if include_init else []
        u = u0
        for _ in range(n):
            u = stepper_fn(u)
            out.append(u)
        return np.stack(out)
    return run


def repeat(stepper_fn, n):
    """``_utils.py:189-254`` (takes_aux=False): final state after n steps."""
    def run(u0):
        u = u0
        for _ in range(n):
            u = stepper_fn(u)
        return u
    return run


def nrmse(u_pred, u_ref):
    """``metrics/_spatial.py:537-586`` with ``spatial_norm(mode="normalized")``:
    sum over channels of ||u - u_ref||_2 / ||u_ref||_2 (the (L/N)^D factor
    cancels). Computed in float64. Works on (C, ...) states."""
    a = np.asarray(u_pred, dtype=np.float64)
    b = np.asarray(u_ref, dtype=np.float64)
    c = a.shape[0]
    num = np.sqrt(np.sum((a - b).reshape(c, -1) ** 2, axis=1))
    den = np.sqrt(np.sum(b.reshape(c, -1) ** 2, axis=1))
    return float(np.sum(num / den))


class Ensemble:
    """Members that differ only in their linear operator, advanced together:
    the vmap over stepper constructors of the reference
    (docs/examples/performance_hints.ipynb cells 23-31). The members' linear
    operators are stacked to (B, C, ...) and broadcast against (B, C, ...)
    states; the nonlinear function is shared."""

    def __init__(self, members, *, dtype=np.float64, coef_dtype=np.float64, order=2, num_circle_points=16,
                 circle_radius=1.0):
        m0 = members[0]
        self.num_spatial_dims = m0.num_spatial_dims
        self.num_points = m0.num_points
        self.num_channels = m0.num_channels
        self.dtype = np.dtype(dtype)
        lin = np.stack([m.linear_operator for m in members])
        kw = dict(dtype=dtype, coef_dtype=coef_dtype)
        ckw = dict(num_circle_points=num_circle_points, circle_radius=circle_radius)
        cls = {0: et.ETDRK0, 1: et.ETDRK1, 2: et.ETDRK2, 3: et.ETDRK3, 4: et.ETDRK4}[order]
        if order == 0:
            self._integrator = cls(m0.dt, lin, **kw)
        else:
            self._integrator = cls(m0.dt, lin, m0.nonlinear_fun, **ckw, **kw)

    def step_batch(self, ub):
        ub = np.asarray(ub, dtype=self.dtype)
        uh = sp.fft(ub, num_spatial_dims=self.num_spatial_dims).astype(sp.complex_of(self.dtype))
        nxt = self._integrator.step_fourier(uh)
        return sp.ifft(nxt, num_spatial_dims=self.num_spatial_dims, num_points=self.num_points).astype(
            self.dtype, copy=False)
