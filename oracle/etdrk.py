"""CPU restatement of exponax's ETDRK integrators (test infrastructure only).

Follows ``exponax/etdrk/`` of exponax 0.2.0. Coefficients are formed by the
Kassam-Trefethen contour mean exactly as the reference's ``lax.scan`` bodies,
in ``coef_dtype`` precision (float64 by default; float32 reproduces the
reference's default-precision formation), then cast to the working dtype.
"""
from __future__ import annotations

import numpy as np

from . import spectral as sp


def roots_of_unity(M: int, dtype=np.complex128):
    """``etdrk/_utils.py:9-23``: exp(2 pi i (j - 1/2) / M), j = 1..M."""
    j = np.arange(1, M + 1)
    return np.exp(2j * np.pi * (j - 0.5) / M).astype(dtype)


def _contour_means(linear_operator, dt, M, r, bodies, coef_dtype):
    """Mean over the contour of each body(lr) (real part), in coef_dtype.

    ``_etdrk_2.py:71-87`` / ``_etdrk_3.py:159-185`` / ``_etdrk_4.py:173-191``:
    accumulation order over the roots is sequential as in the scan.
    """
    cdt = sp.complex_of(coef_dtype)
    rdt = np.dtype(coef_dtype)
    L_dt = (np.asarray(linear_operator) * dt).astype(cdt)
    roots = roots_of_unity(M, cdt)
    sums = [np.zeros(L_dt.shape, dtype=rdt) for _ in bodies]
    radius = rdt.type(r)
    for root in roots:
        lr = radius * root + L_dt
        for i, body in enumerate(bodies):
            sums[i] = sums[i] + body(lr).real.astype(rdt)
    return [dt * (s / rdt.type(M)) for s in sums]


class BaseETDRK:
    """``etdrk/_base_etdrk.py:15-80``: E = exp(dt L)."""

    def __init__(self, dt, linear_operator, *, dtype=np.float64, coef_dtype=np.float64):
        self.dt = dt
        self.dtype = np.dtype(dtype)
        self.cdtype = sp.complex_of(dtype)
        self.coef_dtype = np.dtype(coef_dtype)
        cdt = sp.complex_of(coef_dtype)
        L = np.asarray(linear_operator).astype(cdt)
        self._exp_term = np.exp(self.coef_dtype.type(dt) * L).astype(self.cdtype)

    def _cast(self, a):
        a = np.asarray(a)
        return a.astype(self.cdtype if np.iscomplexobj(a) else self.dtype)


class ETDRK0(BaseETDRK):
    """``etdrk/_etdrk_0.py:6-34``."""

    def __init__(self, dt, linear_operator, **kw):
        super().__init__(dt, linear_operator, **kw)

    def step_fourier(self, u_hat):
        return self._exp_term * u_hat


class ETDRK1(BaseETDRK):
    """``etdrk/_etdrk_1.py:10-82``."""

    def __init__(self, dt, linear_operator, nonlinear_fun, *, num_circle_points=16, circle_radius=1.0, **kw):
        super().__init__(dt, linear_operator, **kw)
        self._nonlinear_fun = nonlinear_fun
        (c1,) = _contour_means(linear_operator, dt, num_circle_points, circle_radius,
                               [lambda lr: (np.exp(lr) - 1) / lr], self.coef_dtype)
        self._coef_1 = self._cast(c1)

    def step_fourier(self, u_hat):
        return self._exp_term * u_hat + self._coef_1 * self._nonlinear_fun(u_hat)


class ETDRK2(BaseETDRK):
    """``etdrk/_etdrk_2.py:10-102``."""

    def __init__(self, dt, linear_operator, nonlinear_fun, *, num_circle_points=16, circle_radius=1.0, **kw):
        super().__init__(dt, linear_operator, **kw)
        self._nonlinear_fun = nonlinear_fun
        c1, c2 = _contour_means(
            linear_operator, dt, num_circle_points, circle_radius,
            [lambda lr: (np.exp(lr) - 1) / lr, lambda lr: (np.exp(lr) - 1 - lr) / lr**2],
            self.coef_dtype)
        self._coef_1 = self._cast(c1)
        self._coef_2 = self._cast(c2)

    def step_fourier(self, u_hat):
        n0 = self._nonlinear_fun(u_hat)
        s1 = self._exp_term * u_hat + self._coef_1 * n0
        n1 = self._nonlinear_fun(s1)
        return s1 + self._coef_2 * (n1 - n0)


class ETDRK3(BaseETDRK):
    """``etdrk/_etdrk_3.py:10-212``."""

    def __init__(self, dt, linear_operator, nonlinear_fun, *, num_circle_points=16, circle_radius=1.0, **kw):
        super().__init__(dt, linear_operator, **kw)
        self._nonlinear_fun = nonlinear_fun
        cdt = sp.complex_of(self.coef_dtype)
        L = np.asarray(linear_operator).astype(cdt)
        self._half_exp_term = np.exp(self.coef_dtype.type(0.5 * dt) * L).astype(self.cdtype)
        bodies = [
            lambda lr: (np.exp(lr / 2) - 1) / lr,
            lambda lr: (np.exp(lr) - 1) / lr,
            lambda lr: (-4 - lr + np.exp(lr) * (4 - 3 * lr + lr**2)) / lr**3,
            lambda lr: (4.0 * (2.0 + lr + np.exp(lr) * (-2 + lr))) / lr**3,
            lambda lr: (-4 - 3 * lr - lr**2 + np.exp(lr) * (4 - lr)) / lr**3,
        ]
        cs = _contour_means(linear_operator, dt, num_circle_points, circle_radius, bodies, self.coef_dtype)
        self._coef_1, self._coef_2, self._coef_3, self._coef_4, self._coef_5 = [self._cast(c) for c in cs]

    def step_fourier(self, u_hat):
        n0 = self._nonlinear_fun(u_hat)
        s1 = self._half_exp_term * u_hat + self._coef_1 * n0
        n1 = self._nonlinear_fun(s1)
        s2 = self._exp_term * u_hat + self._coef_2 * (2 * n1 - n0)
        n2 = self._nonlinear_fun(s2)
        return self._exp_term * u_hat + self._coef_3 * n0 + self._coef_4 * n1 + self._coef_5 * n2


class ETDRK4(BaseETDRK):
    """``etdrk/_etdrk_4.py:10-224`` (note coef_5 * 2 * (N1 + N2) at ``:220``)."""

    def __init__(self, dt, linear_operator, nonlinear_fun, *, num_circle_points=16, circle_radius=1.0, **kw):
        super().__init__(dt, linear_operator, **kw)
        self._nonlinear_fun = nonlinear_fun
        cdt = sp.complex_of(self.coef_dtype)
        L = np.asarray(linear_operator).astype(cdt)
        self._half_exp_term = np.exp(self.coef_dtype.type(0.5 * dt) * L).astype(self.cdtype)
        bodies = [
            lambda lr: (np.exp(lr / 2) - 1) / lr,
            lambda lr: (-4 - lr + np.exp(lr) * (4 - 3 * lr + lr**2)) / lr**3,
            lambda lr: (2 + lr + np.exp(lr) * (-2 + lr)) / lr**3,
            lambda lr: (-4 - 3 * lr - lr**2 + np.exp(lr) * (4 - lr)) / lr**3,
        ]
        c1, c4, c5, c6 = _contour_means(linear_operator, dt, num_circle_points, circle_radius, bodies,
                                        self.coef_dtype)
        self._coef_1 = self._cast(c1)
        self._coef_2 = self._coef_1
        self._coef_3 = self._coef_1
        self._coef_4 = self._cast(c4)
        self._coef_5 = self._cast(c5)
        self._coef_6 = self._cast(c6)

    def step_fourier(self, u_hat):
        n0 = self._nonlinear_fun(u_hat)
        s1 = self._half_exp_term * u_hat + self._coef_1 * n0
        n1 = self._nonlinear_fun(s1)
        s2 = self._half_exp_term * u_hat + self._coef_2 * n1
        n2 = self._nonlinear_fun(s2)
        s3 = self._half_exp_term * s1 + self._coef_3 * (2 * n2 - n0)
        n3 = self._nonlinear_fun(s3)
        return (self._exp_term * u_hat + self._coef_4 * n0 + self._coef_5 * 2 * (n1 + n2)
                + self._coef_6 * n3)
