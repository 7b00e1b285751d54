"""Derivatives of the ETDRK step: reverse mode (VJP) and forward mode (JVP).
TEST INFRASTRUCTURE ONLY.

The reference differentiates with ``jax.grad`` through ``lax.scan``
(``_utils.py:251``). JAX is absent here, so this module restates the
derivative by hand, in float64, from the forward definitions in
``oracle/nonlin.py`` and ``oracle/etdrk.py``.

Inner products:
- physical fields use sum(f * g);
- spectra use the Parseval form <a, b> = sum(irfftn(a) * irfftn(b)).

Under these, rfftn and irfftn are mutual adjoints, a diagonal multiplier is
adjoint to its complex conjugate, the dealias mask is self-adjoint, and
removing the mean is self-adjoint. The gradient of a loss l(u) is then
irfftn(VJP_spectral(rfftn(dl/du))).

Pins, in tests/test_oracle_adjoint.py:
- JVP: every nonlinear function here is a polynomial of degree <= 3 in
  u_hat (plus a constant: the Kolmogorov forcing, the Gray-Scott feed). The
  central difference D(h) = (N(u + h d) - N(u - h d)) / (2h) of such a
  function is J d + h^2 T with T independent of h, so (4 D(1/2) - D(1)) / 3
  is its exact derivative along d up to rounding. ``nonlin_jvp`` uses that,
  and the step JVP composes it.
- VJP: pinned by the dot-product identity <J v, w> = <v, J^T w> against that
  JVP, per nonlinear function and per step.
- End to end: the gradient of MSE(u_n, target) is checked against float64
  central finite differences.
"""
from __future__ import annotations

import numpy as np

from . import etdrk as et
from . import nonlin as nl
from . import spectral as sp


def inner(a_hat, b_hat, num_spatial_dims, num_points):
    """Parseval inner product of two half spectra of real fields."""
    return float(np.sum(sp.ifft(a_hat, num_spatial_dims=num_spatial_dims, num_points=num_points)
                        * sp.ifft(b_hat, num_spatial_dims=num_spatial_dims, num_points=num_points)))


# ------------------------------------------------------------ nonlinear terms
def nonlin_jvp(fun, u_hat, d_hat):
    """J(u_hat) d_hat, exact for nonlinear functions of degree <= 3."""
    d1 = 0.5 * (fun(u_hat + d_hat) - fun(u_hat - d_hat))
    d2 = fun(u_hat + 0.5 * d_hat) - fun(u_hat - 0.5 * d_hat)
    return (4.0 * d2 - d1) / 3.0


def _D(fun):
    return fun.derivative_operator.astype(np.complex128)


def _mean_free(f, D):
    axes = tuple(range(-D, 0))
    return f - np.mean(f, axis=axes, keepdims=True)


def nonlin_vjp(fun, u_hat, m_hat):
    """J(u_hat)^T m_hat for the nonlinear function ``fun``, in float64.

    Each branch differentiates the forward formula of the same class in
    oracle/nonlin.py, which follows the reference file cited there.
    """
    D = fun.num_spatial_dims
    if isinstance(fun, nl.ZeroNonlinearFun):
        return np.zeros_like(m_hat)
    if isinstance(fun, nl.GeneralNonlinearFun):
        return (nonlin_vjp(fun.square, u_hat, m_hat) + nonlin_vjp(fun.convection, u_hat, m_hat)
                + nonlin_vjp(fun.gradient_norm, u_hat, m_hat))
    if isinstance(fun, nl.VorticityConvection2d):  # includes the Kolmogorov variant (constant forcing)
        d = _D(fun)
        inv = fun.inv_laplacian.astype(np.complex128)
        uu = fun.ifft(d[1:2] * inv * u_hat)
        vv = fun.ifft(-d[0:1] * inv * u_hat)
        wx = fun.ifft(d[0:1] * u_hat)
        wy = fun.ifft(d[1:2] * u_hat)
        h = -fun.convection_scale * fun.ifft(m_hat)
        return (np.conj(d[1:2] * inv) * fun.fft(wx * h) + np.conj(-d[0:1] * inv) * fun.fft(wy * h)
                + np.conj(d[0:1]) * fun.fft(uu * h) + np.conj(d[1:2]) * fun.fft(vv * h))
    if isinstance(fun, nl.ProjectedConvection3d):  # includes the Kolmogorov variant (constant forcing)
        # N = P fft(u x w), u = ifft(u_hat), w = ifft(dop x u_hat), dop = i kappa.
        # The Leray projector P is real and symmetric per mode, so P^T = P;
        # the adjoint of fft (mask, then rfftn) is ifft (mask, then irfftn)
        # and vice versa; d(u x w) = du x w + u x dw gives the cotangents
        # w x h of u and h x u of w; the per-mode cross product with dop has
        # entries e_{cja} dop_j, whose conjugate transpose is itself (dop is
        # imaginary), so its adjoint is the same cross product.
        ax = u_hat.ndim - 4
        dop = fun.derivative_operator if ax == 0 else fun.derivative_operator[None]
        dop = dop.astype(np.complex128)
        u = fun.ifft(u_hat)
        w = fun.ifft(nl._cross3(dop, u_hat, ax))
        h = fun.ifft(fun.leray(m_hat))
        return fun.fft(nl._cross3(w, h, ax)) + nl._cross3(dop, fun.fft(nl._cross3(h, u, ax)), ax)
    if isinstance(fun, nl.CahnHilliardNonlinearFun):
        # N = s Lap fft(u_0^3), u = ifft(dealias(u_hat)): the adjoint of fft
        # is ifft and that of ifft(dealias(.)) is fft (the mask is binary);
        # Lap is real
        ax = u_hat.ndim - D - 1
        u0 = np.take(fun.ifft(fun.dealias(u_hat)), [0], axis=ax)
        h = fun.ifft(fun.scale * fun.laplace_operator.astype(np.complex128) * m_hat)
        g = np.zeros(u_hat.shape[:ax] + (u_hat.shape[ax],) + u0.shape[ax + 1:])
        idx = [slice(None)] * g.ndim
        idx[ax] = slice(0, 1)
        g[tuple(idx)] = 3.0 * u0**2 * h
        return fun.fft(g)
    if isinstance(fun, nl.GrayScottNonlinearFun):
        # N0 = f (1 - u0) - u0 u1^2, N1 = -(f + k) u1 + u0 u1^2
        ax = u_hat.ndim - D - 1
        u = fun.ifft(fun.dealias(u_hat))
        h = fun.ifft(m_hat)
        u0, u1 = np.take(u, 0, axis=ax), np.take(u, 1, axis=ax)
        h0, h1 = np.take(h, 0, axis=ax), np.take(h, 1, axis=ax)
        f, k = fun.feed_rate, fun.kill_rate
        return fun.fft(np.stack([(h1 - h0) * u1**2 - f * h0, 2.0 * u0 * u1 * (h1 - h0) - (f + k) * h1], axis=ax))
    if isinstance(fun, nl.BelousovZhabotinskyNonlinearFun):
        # N0 = u0 + u1 - u0 u1 - u0^2, N1 = u2 - u1 - u0 u1, N2 = u0 - u2
        ax = u_hat.ndim - D - 1
        u = fun.ifft(fun.dealias(u_hat))
        h = fun.ifft(m_hat)
        u0, u1 = np.take(u, 0, axis=ax), np.take(u, 1, axis=ax)
        h0, h1, h2 = (np.take(h, i, axis=ax) for i in range(3))
        return fun.fft(np.stack([h0 * (1.0 - u1 - 2.0 * u0) - h1 * u1 + h2, h0 * (1.0 - u0) - h1 * (1.0 + u0),
                                 h1 - h2], axis=ax))
    if isinstance(fun, nl.GradientNormNonlinearFun):
        d = _D(fun)
        grad = fun.ifft(d[None, :] * np.expand_dims(u_hat, -(D + 1)))
        h = -fun.scale * 0.5 * fun.ifft(m_hat)
        if fun.zero_mode_fix:
            h = _mean_free(h, D)
        return np.sum(np.conj(d) * fun.fft(2.0 * grad * np.expand_dims(h, -(D + 1))), axis=-(D + 1))
    if isinstance(fun, nl.PolynomialNonlinearFun):
        u = fun.ifft(u_hat)
        dp = np.zeros_like(u)
        power = np.ones_like(u)
        for k, c in enumerate(fun.coefficients[1:], start=1):
            dp = dp + k * c * power
            power = power * u
        return fun.fft(dp * fun.ifft(m_hat))
    if isinstance(fun, nl.ConvectionNonlinearFun):
        d = _D(fun)
        b = fun.scale
        if fun.single_channel:
            u = fun.ifft(u_hat)
            if fun.conservative:
                # N = -b/2 (sum_d D_d) P F[u^2]
                op = np.sum(d, axis=0, keepdims=True)
                h = fun.ifft(-0.5 * b * np.conj(op) * m_hat)
                return fun.fft(2.0 * u * h)
            # N = -b P F[u * sum_d u_d], u_d = F^-1[P D_d u_hat]
            nab = fun.ifft(d * u_hat)
            s = np.sum(nab, axis=-(D + 1), keepdims=True)
            h = -b * fun.ifft(m_hat)
            return fun.fft(s * h) + np.sum(np.conj(d) * fun.fft(u * h), axis=-(D + 1), keepdims=True)
        u = fun.ifft(u_hat)  # (..., C=D, N...)
        if fun.conservative:
            # N_i = -b/2 sum_j D_j P F[u_i u_j]
            # d/du: cot(u_i u_j) = h_ij = F^-1[P (-b/2) conj(D_j) m_i]
            h = fun.ifft(-0.5 * b * np.conj(d)[None, :] * np.expand_dims(m_hat, -(D + 1)))
            # h[..., i, j, x]; u_i u_j depends on u_k for k = i and k = j
            gu = np.sum(h * np.expand_dims(u, -(D + 2)), axis=-(D + 1)) \
                + np.sum(h * np.expand_dims(u, -(D + 1)), axis=-(D + 2))
            return fun.fft(gu)
        # N_i = -b P F[sum_j u_j d_j u_i], nab[..., i, j] = F^-1[P D_j u_hat_i]
        nab = fun.ifft(d[None, :] * np.expand_dims(u_hat, -(D + 1)))
        h = -b * fun.ifft(m_hat)  # (..., i, x)
        g_u = np.sum(nab * np.expand_dims(h, -(D + 1)), axis=-(D + 2))  # over i: cot u_j
        g_nab = np.expand_dims(u, -(D + 2)) * np.expand_dims(h, -(D + 1))  # [i, j] = u_j h_i
        return fun.fft(g_u) + np.sum(np.conj(d)[None, :] * fun.fft(g_nab), axis=-(D + 1))
    raise NotImplementedError(type(fun).__name__)


# ------------------------------------------------------------------ steps
def step_fourier_jvp(integ, u_hat, d_hat):
    """Derivative of the ETDRK0-4 step_fourier along d_hat
    (etdrk/_etdrk_0.py:33, _etdrk_1.py:78-82, _etdrk_2.py:91-102,
    _etdrk_3.py:191-212, _etdrk_4.py:198-224; the stage names follow
    oracle/etdrk.py)."""
    E = integ._exp_term
    if isinstance(integ, et.ETDRK0):
        return E * d_hat
    f = integ._nonlinear_fun
    jv = lambda w, v: nonlin_jvp(f, w, v)  # noqa: E731
    if isinstance(integ, et.ETDRK1):
        return E * d_hat + integ._coef_1 * jv(u_hat, d_hat)
    if isinstance(integ, et.ETDRK2):
        c1, c2 = integ._coef_1, integ._coef_2
        n0 = f(u_hat)
        s1 = E * u_hat + c1 * n0
        dn0 = jv(u_hat, d_hat)
        ds1 = E * d_hat + c1 * dn0
        dn1 = jv(s1, ds1)
        return ds1 + c2 * (dn1 - dn0)
    Eh = integ._half_exp_term
    if isinstance(integ, et.ETDRK3):
        c1, c2, c3, c4, c5 = integ._coef_1, integ._coef_2, integ._coef_3, integ._coef_4, integ._coef_5
        n0 = f(u_hat)
        ua = Eh * u_hat + c1 * n0
        n1 = f(ua)
        ub = E * u_hat + c2 * (2 * n1 - n0)
        dn0 = jv(u_hat, d_hat)
        dn1 = jv(ua, Eh * d_hat + c1 * dn0)
        dn2 = jv(ub, E * d_hat + c2 * (2 * dn1 - dn0))
        return E * d_hat + c3 * dn0 + c4 * dn1 + c5 * dn2
    if isinstance(integ, et.ETDRK4):
        c1, c4, c5, c6 = integ._coef_1, integ._coef_4, integ._coef_5, integ._coef_6
        n0 = f(u_hat)
        ua = Eh * u_hat + c1 * n0
        n1 = f(ua)
        ub = Eh * u_hat + c1 * n1
        n2 = f(ub)
        uc = Eh * ua + c1 * (2 * n2 - n0)
        dn0 = jv(u_hat, d_hat)
        dua = Eh * d_hat + c1 * dn0
        dn1 = jv(ua, dua)
        dn2 = jv(ub, Eh * d_hat + c1 * dn1)
        dn3 = jv(uc, Eh * dua + c1 * (2 * dn2 - dn0))
        return E * d_hat + c4 * dn0 + 2 * c5 * (dn1 + dn2) + c6 * dn3
    raise NotImplementedError(type(integ).__name__)


def step_fourier_vjp(integ, u_hat, lam):
    """Transpose of ``step_fourier_jvp`` applied to the cotangent lam.

    ETDRK2, with N0 = N(u), S1 = E u + c1 N0, u' = S1 + c2 (N(S1) - N0):
        rho = lam + J(S1)^T (c2 lam)
        out = E rho + J(u)^T (c1 rho - c2 lam)
    ETDRK3 (ua = Eh u + c1 N0, ub = E u + c2 (2 N(ua) - N0),
    u' = E u + c3 N0 + c4 N(ua) + c5 N(ub)):
        l_ub = J(ub)^T (c5 lam)
        l_ua = J(ua)^T (c4 lam + 2 c2 l_ub)
        out = E (lam + l_ub) + Eh l_ua + J(u)^T (c3 lam - c2 l_ub + c1 l_ua)
    ETDRK4 (ua = Eh u + c1 N0, ub = Eh u + c1 N(ua),
    uc = Eh ua + c1 (2 N(ub) - N0),
    u' = E u + c4 N0 + 2 c5 (N(ua) + N(ub)) + c6 N(uc)):
        l_uc = J(uc)^T (c6 lam)
        l_ub = J(ub)^T (2 c5 lam + 2 c1 l_uc)
        l_ua = Eh l_uc + J(ua)^T (2 c5 lam + c1 l_ub)
        out = E lam + Eh (l_ub + l_ua) + J(u)^T (c4 lam - c1 l_uc + c1 l_ua)
    Every multiplier is conjugated: the coefficients are real for real
    linear operators, and conj() keeps the formulas right for complex ones.
    """
    E = np.conj(integ._exp_term)
    if isinstance(integ, et.ETDRK0):
        return E * lam
    f = integ._nonlinear_fun
    vj = lambda w, m: nonlin_vjp(f, w, m)  # noqa: E731
    if isinstance(integ, et.ETDRK1):
        return E * lam + vj(u_hat, np.conj(integ._coef_1) * lam)
    if isinstance(integ, et.ETDRK2):
        c1, c2 = np.conj(integ._coef_1), np.conj(integ._coef_2)
        s1 = integ._exp_term * u_hat + integ._coef_1 * f(u_hat)
        rho = lam + vj(s1, c2 * lam)
        return E * rho + vj(u_hat, c1 * rho - c2 * lam)
    Eh = np.conj(integ._half_exp_term)
    if isinstance(integ, et.ETDRK3):
        n0 = f(u_hat)
        ua = integ._half_exp_term * u_hat + integ._coef_1 * n0
        ub = integ._exp_term * u_hat + integ._coef_2 * (2 * f(ua) - n0)
        c1, c2, c3, c4, c5 = (np.conj(c) for c in (integ._coef_1, integ._coef_2, integ._coef_3, integ._coef_4,
                                                      integ._coef_5))
        l_ub = vj(ub, c5 * lam)
        l_ua = vj(ua, c4 * lam + 2 * c2 * l_ub)
        return E * (lam + l_ub) + Eh * l_ua + vj(u_hat, c3 * lam - c2 * l_ub + c1 * l_ua)
    if isinstance(integ, et.ETDRK4):
        n0 = f(u_hat)
        ua = integ._half_exp_term * u_hat + integ._coef_1 * n0
        ub = integ._half_exp_term * u_hat + integ._coef_1 * f(ua)
        uc = integ._half_exp_term * ua + integ._coef_1 * (2 * f(ub) - n0)
        c1, c4, c5, c6 = (np.conj(c) for c in (integ._coef_1, integ._coef_4, integ._coef_5, integ._coef_6))
        l_uc = vj(uc, c6 * lam)
        l_ub = vj(ub, 2 * c5 * lam + 2 * c1 * l_uc)
        l_ua = Eh * l_uc + vj(ua, 2 * c5 * lam + c1 * l_ub)
        return E * lam + Eh * (l_ub + l_ua) + vj(u_hat, c4 * lam - c1 * l_uc + c1 * l_ua)
    raise NotImplementedError(type(integ).__name__)


def _integrator(stepper):
    return stepper._integrator


def step_jvp(stepper, u, v):
    """d step(u) / du along v, physical space (float64)."""
    D = stepper.num_spatial_dims
    uh = sp.fft(np.asarray(u, np.float64), num_spatial_dims=D)
    vh = sp.fft(np.asarray(v, np.float64), num_spatial_dims=D)
    return sp.ifft(step_fourier_jvp(_integrator(stepper), uh, vh), num_spatial_dims=D,
                   num_points=stepper.num_points)


def step_vjp(stepper, u, g):
    """(d step(u) / du)^T g, physical space (float64)."""
    D = stepper.num_spatial_dims
    uh = sp.fft(np.asarray(u, np.float64), num_spatial_dims=D)
    gh = sp.fft(np.asarray(g, np.float64), num_spatial_dims=D)
    return sp.ifft(step_fourier_vjp(_integrator(stepper), uh, gh), num_spatial_dims=D,
                   num_points=stepper.num_points)


def repeat_vjp(stepper, u0, n, g_final):
    """Gradient through ``repeat(stepper, n)`` (_utils.py:189-254): all
    intermediate states are kept (float64, small cases)."""
    step = stepper.step_batch
    states = [np.asarray(u0, np.float64)]
    for _ in range(n - 1):
        states.append(step(states[-1]))
    g = np.asarray(g_final, np.float64)
    for u in reversed(states):
        g = step_vjp(stepper, u, g)
    return g


def rollout_vjp(stepper, u0, n, g_traj, *, include_init=False):
    """Gradient through ``rollout(stepper, n, include_init)`` (_utils.py:92-186)
    for a cotangent g_traj of the stacked trajectory (leading time axis)."""
    step = stepper.step_batch
    states = [np.asarray(u0, np.float64)]
    for _ in range(n - 1):
        states.append(step(states[-1]))
    off = 1 if include_init else 0
    g = np.zeros_like(states[0])
    for t in range(n - 1, -1, -1):
        g = g + np.asarray(g_traj[t + off], np.float64)
        g = step_vjp(stepper, states[t], g)
    if include_init:
        g = g + np.asarray(g_traj[0], np.float64)
    return g


def mse_grad(stepper, u0, target, n):
    """Loss mean((repeat(stepper, n)(u0) - target)^2) and its gradient in u0."""
    step = stepper.step_batch
    u = np.asarray(u0, np.float64)
    for _ in range(n):
        u = step(u)
    r = u - target
    loss = float(np.mean(r * r))
    return loss, repeat_vjp(stepper, u0, n, 2.0 * r / r.size)
