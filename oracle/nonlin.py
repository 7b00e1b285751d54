"""CPU restatement of exponax's nonlinear functions (test infrastructure only).

Follows ``exponax/nonlin_fun/`` of exponax 0.2.0 (file:line per class). All
operators are cast to the working precision given by ``dtype`` so that the
FP32 oracle computes in float32/complex64 like the reference's default JAX
precision, and the FP64 oracle in float64/complex128.
"""
from __future__ import annotations

import numpy as np

from . import spectral as sp


class BaseNonlinearFun:
    """``nonlin_fun/_base.py:9-159``: dealiased fft/ifft wrappers."""

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction=None, dtype=np.float64):
        self.num_spatial_dims = num_spatial_dims
        self.num_points = num_points
        self.dtype = np.dtype(dtype)
        self.cdtype = sp.complex_of(dtype)
        self.dealiasing_mask = sp.dealiasing_mask(num_spatial_dims, num_points, dealiasing_fraction)

    def dealias(self, u_hat):
        """``_base.py:73-97``."""
        if self.dealiasing_mask is None:
            raise ValueError("Nonlinear function was set up without dealiasing")
        return self.dealiasing_mask * u_hat

    def fft(self, u):
        """``_base.py:99-116``: forward transform then post-dealias."""
        u_hat = sp.fft(u.astype(self.dtype, copy=False), num_spatial_dims=self.num_spatial_dims)
        if self.dealiasing_mask is not None:
            u_hat = self.dealiasing_mask * u_hat
        return u_hat.astype(self.cdtype, copy=False)

    def ifft(self, u_hat):
        """``_base.py:118-137``: pre-dealias then inverse transform."""
        if self.dealiasing_mask is not None:
            u_hat = self.dealiasing_mask * u_hat
        u = sp.ifft(u_hat.astype(self.cdtype, copy=False), num_spatial_dims=self.num_spatial_dims,
                    num_points=self.num_points)
        return u.astype(self.dtype, copy=False)

    def _op(self, a):
        return np.asarray(a).astype(self.cdtype)


class ZeroNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_zero.py:7-38``."""

    def __init__(self, num_spatial_dims, num_points, *, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, dtype=dtype)

    def __call__(self, u_hat):
        return np.zeros_like(u_hat)


class ConvectionNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_convection.py:7-245`` (four variants, all times -scale)."""

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction=2 / 3,
                 scale=1.0, single_channel=False, conservative=False, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.derivative_operator = self._op(derivative_operator)
        self.scale = scale
        self.single_channel = single_channel
        self.conservative = conservative

    def _multi_channel_conservative(self, u_hat):
        """``_convection.py:104-138``: -b/2 sum_j i kappa_j F[u_i u_j]."""
        D = self.num_spatial_dims
        if u_hat.shape[-(D + 1)] != D:
            raise ValueError("Number of channels in u_hat should match number of spatial dimensions")
        u = self.ifft(u_hat)
        outer = np.expand_dims(u, -(D + 2)) * np.expand_dims(u, -(D + 1))
        outer_hat = self.fft(outer)
        conv = 0.5 * np.sum(self.derivative_operator[None, :] * outer_hat, axis=-(D + 1))
        return (-self.scale * conv).astype(self.cdtype, copy=False)

    def _multi_channel_nonconservative(self, u_hat):
        """``_convection.py:140-171``: -b F[sum_j u_j d_j u_i]."""
        D = self.num_spatial_dims
        if u_hat.shape[-(D + 1)] != D:
            raise ValueError("Number of channels in u_hat should match number of spatial dimensions")
        u = self.ifft(u_hat)
        nabla_u = self.ifft(self.derivative_operator[None, :] * np.expand_dims(u_hat, -(D + 1)))
        conv = np.sum(np.expand_dims(u, -(D + 2)) * nabla_u, axis=-(D + 1))
        return (-self.scale * self.fft(conv)).astype(self.cdtype, copy=False)

    def _single_channel_conservative(self, u_hat):
        """``_convection.py:173-202``: -b/2 (sum_d i kappa_d) F[u^2]."""
        u = self.ifft(u_hat)
        sq_hat = self.fft(u**2)
        op = np.sum(self.derivative_operator, axis=0, keepdims=True)
        return (-self.scale * (0.5 * op * sq_hat)).astype(self.cdtype, copy=False)

    def _single_channel_nonconservative(self, u_hat):
        """``_convection.py:204-231``: -b F[u sum_d d_d u]."""
        u = self.ifft(u_hat)
        D = self.num_spatial_dims
        nabla_u = self.ifft(self.derivative_operator * u_hat)
        conv = np.sum(u * nabla_u, axis=-(D + 1), keepdims=True)
        return (-self.scale * self.fft(conv)).astype(self.cdtype, copy=False)

    def __call__(self, u_hat):
        """``_convection.py:233-245``."""
        if self.single_channel:
            if self.conservative:
                return self._single_channel_conservative(u_hat)
            return self._single_channel_nonconservative(u_hat)
        if self.conservative:
            return self._multi_channel_conservative(u_hat)
        return self._multi_channel_nonconservative(u_hat)


class GradientNormNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_gradient_norm.py:8-101``: -b/2 F[|grad u|^2 - mean]."""

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction,
                 zero_mode_fix=True, scale=1.0, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.derivative_operator = self._op(derivative_operator)
        self.zero_mode_fix = zero_mode_fix
        self.scale = scale

    def __call__(self, u_hat):
        D = self.num_spatial_dims
        grad_hat = self.derivative_operator[None, :] * np.expand_dims(u_hat, -(D + 1))
        grad = self.ifft(grad_hat)
        gn = np.sum(grad**2, axis=-(D + 1))
        if self.zero_mode_fix:
            # ``_gradient_norm.py:78-82,94-96``: per channel f - mean(f)
            axes = tuple(range(-D, 0))
            gn = gn - np.mean(gn, axis=axes, keepdims=True).astype(self.dtype)
        gn_hat = 0.5 * self.fft(gn)
        return (-self.scale * gn_hat).astype(self.cdtype, copy=False)


class PolynomialNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_polynomial.py:6-76``: F[sum_k c_k u^k] (no minus)."""

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction, coefficients, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.coefficients = tuple(coefficients)

    def __call__(self, u_hat):
        u = self.ifft(u_hat)
        power = np.ones_like(u)
        acc = np.zeros_like(u)
        for c in self.coefficients:
            acc = acc + self.dtype.type(c) * power
            power = power * u
        return self.fft(acc)


class CahnHilliardNonlinearFun(BaseNonlinearFun):
    """``stepper/reaction/_cahn_hilliard.py:10-41``: scale * Laplace F[u_0^3]
    (first channel), dealiased before the inverse and after the forward
    transform."""

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, scale, dealiasing_fraction,
                 dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.laplace_operator = self._op(sp.build_laplace_operator(derivative_operator, order=2))
        self.scale = scale

    def __call__(self, u_hat):
        u = self.ifft(self.dealias(u_hat))
        first = np.take(u, [0], axis=u.ndim - self.num_spatial_dims - 1)  # channel axis, batched or not
        power_hat = self.fft(first**3)
        return (self.laplace_operator * power_hat) * self.dtype.type(self.scale)


class GrayScottNonlinearFun(BaseNonlinearFun):
    """``stepper/reaction/_gray_scott.py:9-49``."""

    def __init__(self, num_spatial_dims, num_points, *, dealiasing_fraction, feed_rate, kill_rate, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.feed_rate, self.kill_rate = feed_rate, kill_rate

    def __call__(self, u_hat):
        u = self.ifft(self.dealias(u_hat))
        ax = u.ndim - self.num_spatial_dims - 1
        if u.shape[ax] != 2:
            raise ValueError("num_channels must be 2")
        u0, u1 = np.take(u, 0, axis=ax), np.take(u, 1, axis=ax)
        f, k = self.dtype.type(self.feed_rate), self.dtype.type(self.kill_rate)
        power = np.stack([f * (1 - u0) - u0 * u1**2, -(f + k) * u1 + u0 * u1**2], axis=ax)
        return self.fft(power)


class BelousovZhabotinskyNonlinearFun(BaseNonlinearFun):
    """``stepper/reaction/_belousov_zhabotinsky.py:9-47``."""

    def __call__(self, u_hat):
        u = self.ifft(self.dealias(u_hat))
        ax = u.ndim - self.num_spatial_dims - 1
        if u.shape[ax] != 3:
            raise ValueError("num_channels must be 3")
        u0, u1, u2 = (np.take(u, i, axis=ax) for i in range(3))
        power = np.stack([u0 + u1 - u0 * u1 - u0**2, u2 - u1 - u0 * u1, u0 - u2], axis=ax)
        return self.fft(power)


def _cross3(a, b, ax):
    """Cross product of two 3-vectors stored along axis ax."""
    g = lambda x, i: np.take(x, i, axis=ax)  # noqa: E731
    return np.stack([g(a, 1) * g(b, 2) - g(a, 2) * g(b, 1), g(a, 2) * g(b, 0) - g(a, 0) * g(b, 2),
                     g(a, 0) * g(b, 1) - g(a, 1) * g(b, 0)], axis=ax)


class Leray:
    """``nonlin_fun/_leray.py:8-103``: u_hat + i kappa p_hat with p_hat =
    -Laplace^-1 (i kappa . u_hat), Laplace^-1 = 0 at the mean mode."""

    def __init__(self, derivative_operator, dtype=np.float64):
        lap = sp.build_laplace_operator(derivative_operator, order=2)
        with np.errstate(divide="ignore"):
            self.inv_laplacian = np.where(lap != 0, 1.0 / np.where(lap != 0, lap, 1.0), 0.0).astype(
                sp.complex_of(dtype))
        self.derivative_operator = derivative_operator.astype(sp.complex_of(dtype))

    def __call__(self, u_hat):
        ax = u_hat.ndim - self.derivative_operator.ndim
        div = np.sum(self.derivative_operator * u_hat, axis=ax, keepdims=True)
        return u_hat + self.derivative_operator * (-self.inv_laplacian * div)


class ProjectedConvection3d(BaseNonlinearFun):
    """``nonlin_fun/_projected_convection.py:19-136``: P(u x omega), omega =
    i kappa x u_hat, the cross product in physical space."""

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction=2 / 3,
                 dtype=np.float64):
        if num_spatial_dims != 3:
            raise ValueError("ProjectedConvection3d only supports 3 spatial dimensions.")
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.derivative_operator = self._op(derivative_operator)
        self.leray = Leray(derivative_operator, dtype=dtype)

    def __call__(self, u_hat):
        ax = u_hat.ndim - 4  # the channel axis
        dop = self.derivative_operator if ax == 0 else self.derivative_operator[None]
        curl_hat = _cross3(dop, u_hat, ax)
        curl = self.ifft(curl_hat)
        velocity = self.ifft(u_hat)
        return self.leray(self.fft(_cross3(velocity, curl, ax)))


class ProjectedConvection3dKolmogorov(ProjectedConvection3d):
    """``nonlin_fun/_projected_convection.py:139-226``."""

    def __init__(self, num_spatial_dims, num_points, *, injection_mode=4, injection_scale=1.0, derivative_operator,
                 dealiasing_fraction, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, derivative_operator=derivative_operator,
                         dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        wn = sp.build_wavenumbers(num_spatial_dims, num_points)
        mask = (wn[0] == 0) & (wn[1] == injection_mode) & (wn[2] == 0)
        single = np.where(mask, injection_scale * sp.build_scaling_array(num_spatial_dims, num_points,
                                                                         mode="coef_extraction"), 0.0)
        zeros = np.zeros_like(single)
        self.injection = np.concatenate([single, zeros, zeros], axis=0).astype(self.cdtype)

    def __call__(self, u_hat):
        return super().__call__(u_hat) + self.injection


class GeneralNonlinearFun(BaseNonlinearFun):
    """``nonlin_fun/_general_nonlinear.py:9-119``: b0 u^2 + conv + grad-norm."""

    def __init__(self, num_spatial_dims, num_points, *, derivative_operator, dealiasing_fraction,
                 scale_list=(0.0, -1.0, 0.0), zero_mode_fix=True, dtype=np.float64):
        if len(scale_list) != 3:
            raise ValueError("The scale list must have exactly 3 elements")
        self.square = PolynomialNonlinearFun(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction,
                                             coefficients=[0.0, 0.0, scale_list[0]], dtype=dtype)
        self.convection = ConvectionNonlinearFun(num_spatial_dims, num_points,
                                                 derivative_operator=derivative_operator,
                                                 dealiasing_fraction=dealiasing_fraction, scale=-scale_list[1],
                                                 single_channel=True, conservative=True, dtype=dtype)
        self.gradient_norm = GradientNormNonlinearFun(num_spatial_dims, num_points,
                                                      derivative_operator=derivative_operator,
                                                      dealiasing_fraction=dealiasing_fraction,
                                                      scale=-scale_list[2], zero_mode_fix=zero_mode_fix,
                                                      dtype=dtype)
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)

    def __call__(self, u_hat):
        return self.square(u_hat) + self.convection(u_hat) + self.gradient_norm(u_hat)


class VorticityConvection2d(BaseNonlinearFun):
    """``nonlin_fun/_vorticity_convection.py:8-99``: -b F[u w_x + v w_y]."""

    def __init__(self, num_spatial_dims, num_points, *, convection_scale=1.0, derivative_operator,
                 dealiasing_fraction, dtype=np.float64):
        if num_spatial_dims != 2:
            raise ValueError(f"Expected num_spatial_dims = 2, got {num_spatial_dims}.")
        super().__init__(num_spatial_dims, num_points, dealiasing_fraction=dealiasing_fraction, dtype=dtype)
        self.convection_scale = convection_scale
        d64 = np.asarray(derivative_operator)
        self.derivative_operator = self._op(d64)
        lap = sp.build_laplace_operator(d64, order=2)
        with np.errstate(divide="ignore", invalid="ignore"):  # the mean mode (lap = 0) is replaced by 1
            inv = np.where(lap == 0, 1.0, 1.0 / lap)
        self.inv_laplacian = self._op(inv)

    def __call__(self, u_hat):
        w_hat = u_hat
        psi_hat = self.inv_laplacian * w_hat
        uu_hat = self.derivative_operator[1:2] * psi_hat
        vv_hat = -self.derivative_operator[0:1] * psi_hat
        wx_hat = self.derivative_operator[0:1] * w_hat
        wy_hat = self.derivative_operator[1:2] * w_hat
        uu = self.ifft(uu_hat)
        vv = self.ifft(vv_hat)
        wx = self.ifft(wx_hat)
        wy = self.ifft(wy_hat)
        conv = uu * wx + vv * wy
        return (-self.convection_scale * self.fft(conv)).astype(self.cdtype, copy=False)


class VorticityConvection2dKolmogorov(VorticityConvection2d):
    """``nonlin_fun/_vorticity_convection.py:102-182``: adds the forcing
    -k_f * gamma * scaling(coef_extraction) at wavenumber (0, k_f)."""

    def __init__(self, num_spatial_dims, num_points, *, convection_scale=1.0, injection_mode=4,
                 injection_scale=1.0, derivative_operator, dealiasing_fraction, dtype=np.float64):
        super().__init__(num_spatial_dims, num_points, convection_scale=convection_scale,
                         derivative_operator=derivative_operator, dealiasing_fraction=dealiasing_fraction,
                         dtype=dtype)
        wn = sp.build_wavenumbers(num_spatial_dims, num_points)
        where = (wn[0] == 0) & (wn[1] == injection_mode)
        scal = sp.build_scaling_array(num_spatial_dims, num_points, mode="coef_extraction")
        self.injection = self._op(np.where(where, -injection_mode * injection_scale * scal, 0.0))

    def __call__(self, u_hat):
        return super().__call__(u_hat) + self.injection
