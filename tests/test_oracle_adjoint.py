"""Pins of the oracle's derivative module (oracle/adjoint.py) on the CPU.

- Each nonlinear VJP is checked by the dot-product identity against the exact
  JVP of the quadratic nonlinear function.
- Each step VJP is checked by the dot-product identity against the step JVP.
- The step JVP is checked against central differences.
- The gradient of MSE after a rollout is checked against float64 central
  differences with eps = 1e-3, on random directions. This is the KS 2-D
  gradient-norm case at 128^2, 50 steps.

A planted mistake shows that the dot-product check detects errors.
"""
import numpy as np
import pytest

from oracle import adjoint as ad
from oracle import ic
from oracle import nonlin as nl
from oracle import spectral as sp
from oracle import stepper as st


def _fields(D, N, C, seed, cutoff=4, L=1.0):
    rng = np.random.default_rng(seed)
    return np.stack([ic.tfs(D, L, N, cutoff, rng)[0] for _ in range(C)])


def _nonlin_cases():
    out = []
    for D in (1, 2):
        N = 32 if D == 1 else 16
        d = sp.build_derivative_operator(D, 3.0, N)
        out += [
            (f"gradnorm{D}", D, N, 1, nl.GradientNormNonlinearFun(D, N, derivative_operator=d,
                                                                  dealiasing_fraction=2 / 3, scale=0.7)),
            (f"gradnorm_nofix{D}", D, N, 1, nl.GradientNormNonlinearFun(D, N, derivative_operator=d,
                                                                        dealiasing_fraction=2 / 3,
                                                                        zero_mode_fix=False)),
            (f"conv_sc_cons{D}", D, N, 1, nl.ConvectionNonlinearFun(D, N, derivative_operator=d, scale=1.3,
                                                                    single_channel=True, conservative=True)),
            (f"conv_sc_noncons{D}", D, N, 1, nl.ConvectionNonlinearFun(D, N, derivative_operator=d,
                                                                       single_channel=True)),
            (f"conv_mc_cons{D}", D, N, D, nl.ConvectionNonlinearFun(D, N, derivative_operator=d,
                                                                    conservative=True)),
            (f"conv_mc_noncons{D}", D, N, D, nl.ConvectionNonlinearFun(D, N, derivative_operator=d, scale=0.4)),
            (f"poly{D}", D, N, 1, nl.PolynomialNonlinearFun(D, N, dealiasing_fraction=2 / 3,
                                                            coefficients=(0.3, -1.0, 2.0))),
            (f"general{D}", D, N, 1, nl.GeneralNonlinearFun(D, N, derivative_operator=d, dealiasing_fraction=2 / 3,
                                                            scale_list=(0.5, -1.0, 0.25))),
        ]
    d2 = sp.build_derivative_operator(2, 2 * np.pi, 16)
    out += [
        ("vorticity", 2, 16, 1, nl.VorticityConvection2d(2, 16, convection_scale=0.9, derivative_operator=d2,
                                                         dealiasing_fraction=2 / 3)),
        ("kolmogorov", 2, 16, 1, nl.VorticityConvection2dKolmogorov(2, 16, derivative_operator=d2,
                                                                    dealiasing_fraction=2 / 3)),
    ]
    for D in (1, 2):
        N = 32 if D == 1 else 16
        d = sp.build_derivative_operator(D, 3.0, N)
        out += [
            (f"allen_cahn_cubic{D}", D, N, 1, nl.PolynomialNonlinearFun(D, N, dealiasing_fraction=2 / 3,
                                                                        coefficients=(0.0, 1.0, 0.0, -1.0))),
            (f"cahn_hilliard{D}", D, N, 1, nl.CahnHilliardNonlinearFun(D, N, derivative_operator=d, scale=0.6,
                                                                       dealiasing_fraction=2 / 3)),
            (f"gray_scott{D}", D, N, 2, nl.GrayScottNonlinearFun(D, N, dealiasing_fraction=2 / 3, feed_rate=0.04,
                                                                 kill_rate=0.06)),
            (f"belousov_zhabotinsky{D}", D, N, 3, nl.BelousovZhabotinskyNonlinearFun(D, N,
                                                                                    dealiasing_fraction=2 / 3)),
        ]
    d3 = sp.build_derivative_operator(3, 2 * np.pi, 8)
    out += [
        ("projected3d", 3, 8, 3, nl.ProjectedConvection3d(3, 8, derivative_operator=d3, dealiasing_fraction=2 / 3)),
        ("projected3d_kolmogorov", 3, 8, 3, nl.ProjectedConvection3dKolmogorov(
            3, 8, injection_mode=1, injection_scale=0.7, derivative_operator=d3, dealiasing_fraction=2 / 3)),
    ]
    return out


NONLIN = _nonlin_cases()


@pytest.mark.parametrize("name,D,N,C,fun", NONLIN, ids=[c[0] for c in NONLIN])
def test_nonlinear_vjp_dot_product(name, D, N, C, fun):
    uh = sp.fft(_fields(D, N, C, 1), num_spatial_dims=D)
    vh = sp.fft(_fields(D, N, C, 2, cutoff=7), num_spatial_dims=D)
    wh = sp.fft(_fields(D, N, C, 3, cutoff=7), num_spatial_dims=D)
    jv = ad.nonlin_jvp(fun, uh, vh)
    jtw = ad.nonlin_vjp(fun, uh, wh)
    lhs = ad.inner(jv, wh, D, N)
    rhs = ad.inner(vh, jtw, D, N)
    assert abs(lhs) > 1e-6
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1.0), (lhs, rhs)


def test_nonlinear_jvp_is_exact_derivative():
    """For the quadratic nonlinearity the central difference with any step is
    the derivative: it equals a tiny-step difference."""
    d = sp.build_derivative_operator(1, 5.0, 32)
    f = nl.ConvectionNonlinearFun(1, 32, derivative_operator=d, single_channel=True, conservative=True)
    uh = sp.fft(_fields(1, 32, 1, 4), num_spatial_dims=1)
    vh = sp.fft(_fields(1, 32, 1, 5), num_spatial_dims=1)
    a = ad.nonlin_jvp(f, uh, vh)
    eps = 1e-6
    b = (f(uh + eps * vh) - f(uh - eps * vh)) / (2 * eps)
    np.testing.assert_allclose(a, b, atol=1e-7 * np.abs(a).max())


def test_dot_product_detects_planted_error():
    """Removing the mean removal from the gradient-norm VJP (a plausible slip)
    breaks the identity."""
    d = sp.build_derivative_operator(2, 3.0, 16)
    f = nl.GradientNormNonlinearFun(2, 16, derivative_operator=d, dealiasing_fraction=2 / 3)
    uh = sp.fft(_fields(2, 16, 1, 1), num_spatial_dims=2)
    vh = sp.fft(_fields(2, 16, 1, 2, cutoff=7), num_spatial_dims=2)
    wh = sp.fft(_fields(2, 16, 1, 3, cutoff=7), num_spatial_dims=2)
    wh[..., 0, 0] += 5.0  # a mean in the cotangent makes the slip visible
    lhs = ad.inner(ad.nonlin_jvp(f, uh, vh), wh, 2, 16)
    f_bad = nl.GradientNormNonlinearFun(2, 16, derivative_operator=d, dealiasing_fraction=2 / 3)
    f_bad.zero_mode_fix = False  # the VJP now skips the mean removal the forward did
    rhs_bad = ad.inner(vh, ad.nonlin_vjp(f_bad, uh, wh), 2, 16)
    rhs = ad.inner(vh, ad.nonlin_vjp(f, uh, wh), 2, 16)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    assert abs(lhs - rhs_bad) > 1e-3 * abs(lhs)


def _steppers():
    return [
        ("ks2d", st.KuramotoSivashinsky(2, 20.0, 16, 0.05), 1),
        ("ks_cons1d", st.KuramotoSivashinskyConservative(1, 30.0, 64, 0.1), 1),
        ("burgers1d", st.Burgers(1, 1.0, 64, 1e-3, diffusivity=0.01), 1),
        ("burgers2d", st.Burgers(2, 1.0, 16, 1e-3, diffusivity=0.02), 2),
        ("kolmogorov", st.KolmogorovFlowVorticity(2, 2 * np.pi, 16, 0.01, diffusivity=0.01), 1),
        ("ks1d_order1", st.KuramotoSivashinsky(1, 30.0, 64, 0.02, order=1), 1),
        ("ks1d_order0", st.KuramotoSivashinsky(1, 30.0, 64, 0.02, order=0), 1),
        # complex L: the transpose multiplies by conj(E)
        ("kdv1d", st.KortewegDeVries(1, 20.0, 64, 1e-3), 1),
        ("kdv1d_order1", st.KortewegDeVries(1, 20.0, 64, 1e-3, order=1), 1),
        ("dispersion1d", st.Dispersion(1, 1.0, 64, 1e-3, dispersivity=0.01), 1),
        ("general1d_complex", st.GeneralStepper(1, 2.0, 64, 1e-3, linear_coefficients=(0.0, 0.1, 0.01),
                                                nonlinear_scales=(0.3, -1.0, -0.5)), 1),
        # ETDRK3 and ETDRK4 (etdrk/_etdrk_3.py:191-212, _etdrk_4.py:198-224)
        ("ks1d_order3", st.KuramotoSivashinsky(1, 30.0, 64, 0.05, order=3), 1),
        ("ks1d_order4", st.KuramotoSivashinsky(1, 30.0, 64, 0.05, order=4), 1),
        ("kdv1d_order4", st.KortewegDeVries(1, 20.0, 64, 1e-3, order=4), 1),
        ("burgers2d_order3", st.Burgers(2, 1.0, 16, 1e-3, diffusivity=0.02, order=3), 2),
        ("kolmogorov_order4", st.KolmogorovFlowVorticity(2, 2 * np.pi, 16, 0.01, diffusivity=0.01, order=4), 1),
        ("general2d_order4", st.GeneralStepper(2, 2.0, 16, 1e-3, linear_coefficients=(0.0, 0.1, 0.01),
                                               nonlinear_scales=(0.3, -1.0, -0.5), order=4), 1),
    ]


STEPPERS = _steppers()


@pytest.mark.parametrize("name,s,C", STEPPERS, ids=[c[0] for c in STEPPERS])
def test_step_vjp_dot_product_and_jvp_fd(name, s, C):
    D, N = s.num_spatial_dims, s.num_points
    u = _fields(D, N, C, 11, L=s.domain_extent)
    v = _fields(D, N, C, 12, cutoff=6, L=s.domain_extent)
    w = _fields(D, N, C, 13, cutoff=6, L=s.domain_extent)
    jv = ad.step_jvp(s, u, v)
    jtw = ad.step_vjp(s, u, w)
    lhs, rhs = float(np.sum(jv * w)), float(np.sum(v * jtw))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)
    eps = 1e-4
    fd = (s(u + eps * v) - s(u - eps * v)) / (2 * eps)
    assert np.linalg.norm(fd - jv) <= 1e-7 * np.linalg.norm(jv)


def test_rollout_vjp_matches_repeat_vjp_and_include_init():
    s = st.KuramotoSivashinskyConservative(1, 30.0, 32, 0.1)
    u0 = _fields(1, 32, 1, 21, L=30.0)
    g = _fields(1, 32, 1, 22, L=30.0)
    n = 5
    gt = np.zeros((n, 1, 32))
    gt[-1] = g
    np.testing.assert_allclose(ad.rollout_vjp(s, u0, n, gt), ad.repeat_vjp(s, u0, n, g), atol=1e-12)
    gti = np.concatenate([g[None], gt])
    np.testing.assert_allclose(ad.rollout_vjp(s, u0, n, gti, include_init=True),
                               ad.repeat_vjp(s, u0, n, g) + g, atol=1e-12)


def test_ks2d_mse_gradient_vs_central_fd():
    """KS 2-D (gradient norm), L = 60, N = 128, dt = 0.05, ETDRK2, 50 steps:
    d MSE(u_50, target) / d u0 along random directions against float64
    central differences with eps = 1e-3."""
    s = st.KuramotoSivashinsky(2, 60.0, 128, 0.05)
    rng = np.random.default_rng(4)
    u0 = ic.tfs(2, 60.0, 128, 5, rng)
    target = ic.tfs(2, 60.0, 128, 5, np.random.default_rng(40))
    n = 50
    loss, g = ad.mse_grad(s, u0, target, n)
    assert np.isfinite(loss) and np.isfinite(g).all()

    def loss_of(u):
        for _ in range(n):
            u = s(u)
        return float(np.mean((u - target) ** 2))

    drng = np.random.default_rng(41)
    for _ in range(2):
        v = drng.standard_normal(u0.shape)
        eps = 1e-3
        fd = (loss_of(u0 + eps * v) - loss_of(u0 - eps * v)) / (2 * eps)
        an = float(np.sum(g * v))
        assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
