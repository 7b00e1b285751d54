"""Pins of the oracle's nonlinear functions (oracle/nonlin.py) and ETDRK
integrators (oracle/etdrk.py) on the CPU: the closed forms of the
reference's tests/test_nonlinear_funs.py and tests/test_etdrk.py, the
phi-function closed forms in mpmath, and the convergence orders.
"""
import mpmath
import numpy as np
import pytest

from oracle import etdrk as et
from oracle import nonlin as nl
from oracle import spectral as sp
from oracle import stepper as st


def test_convection_single_channel_sine():
    """tests/test_nonlinear_funs.py:17-52: -u u_x for sin -> -(pi/L) sin(2x')."""
    n, L = 64, 3.0
    d = sp.build_derivative_operator(1, L, n)
    f = nl.ConvectionNonlinearFun(1, n, derivative_operator=d, dealiasing_fraction=2 / 3, scale=1.0,
                                  single_channel=True, conservative=False)
    g = sp.make_grid(1, L, n)
    u = np.sin(2 * np.pi * g / L)
    r = sp.ifft(f(sp.fft(u, num_spatial_dims=1)), num_spatial_dims=1, num_points=n)
    np.testing.assert_allclose(r, -(np.pi / L) * np.sin(2 * 2 * np.pi * g / L), atol=1e-4)


@pytest.mark.parametrize("single,D", [(True, 1), (False, 1), (True, 2), (True, 3)])
def test_conservative_equals_nonconservative(single, D):
    """tests/test_nonlinear_funs.py:55-93 (smooth input); with 2/3
    dealiasing both forms are the same alias-free convolution, so they agree
    to rounding for any band-limited input, which the GPU path relies on.
    In 2-D and 3-D this pins the single-channel non-conservative form
    (_convection.py:204-231) to the conservative one (:173-202)."""
    n, L = {1: 64, 2: 32, 3: 16}[D], 4.0
    d = sp.build_derivative_operator(D, L, n)
    kw = dict(derivative_operator=d, dealiasing_fraction=2 / 3, scale=1.0, single_channel=single)
    cons = nl.ConvectionNonlinearFun(D, n, conservative=True, **kw)
    noncons = nl.ConvectionNonlinearFun(D, n, conservative=False, **kw)
    u = np.random.default_rng(0).standard_normal((1,) + (n,) * D)
    uh = sp.fft(u, num_spatial_dims=D)
    np.testing.assert_allclose(cons(uh), noncons(uh), atol=1e-12 * np.abs(uh).max() ** 2 * n ** D)
    # without dealiasing the forms differ (aliasing), so the check is not vacuous
    c0 = nl.ConvectionNonlinearFun(D, n, conservative=True, derivative_operator=d, dealiasing_fraction=1.0,
                                   single_channel=single)
    n0 = nl.ConvectionNonlinearFun(D, n, conservative=False, derivative_operator=d, dealiasing_fraction=1.0,
                                   single_channel=single)
    assert np.abs(c0(uh) - n0(uh)).max() > 1e-6


def test_single_channel_nonconservative_2d_closed_form():
    """-u (u_x + u_y) for u = sin(2 pi x / L) + sin(2 pi y / L):
    -(2 pi / L) (sin a + sin b)(cos a + cos b)."""
    n, L = 32, 3.0
    d = sp.build_derivative_operator(2, L, n)
    f = nl.ConvectionNonlinearFun(2, n, derivative_operator=d, dealiasing_fraction=2 / 3, scale=1.0,
                                  single_channel=True, conservative=False)
    g = sp.make_grid(2, L, n)
    a, b = 2 * np.pi * g[0] / L, 2 * np.pi * g[1] / L
    u = (np.sin(a) + np.sin(b))[None]
    r = sp.ifft(f(sp.fft(u, num_spatial_dims=2)), num_spatial_dims=2, num_points=n)
    want = -(2 * np.pi / L) * (np.sin(a) + np.sin(b)) * (np.cos(a) + np.cos(b))
    np.testing.assert_allclose(r[0], want, atol=1e-10)


def test_gradient_norm_zero_mode_fix():
    """tests/test_nonlinear_funs.py:96-123."""
    n, L = 64, 3.0
    d = sp.build_derivative_operator(1, L, n)
    f = nl.GradientNormNonlinearFun(1, n, derivative_operator=d, dealiasing_fraction=2 / 3, zero_mode_fix=True)
    g = sp.make_grid(1, L, n)
    uh = sp.fft(np.sin(2 * np.pi * g / L) + 0.5, num_spatial_dims=1)
    assert abs(f(uh)[0, 0]) < 1e-5
    # closed form: -(1/2)((2 pi/L) cos)^2 minus its mean = -(pi/L)^2 cos(2x')
    r = sp.ifft(f(uh), num_spatial_dims=1, num_points=n)
    np.testing.assert_allclose(r, -((2 * np.pi / L) ** 2) * 0.25 * np.cos(2 * 2 * np.pi * g / L), atol=1e-10)


def test_polynomial_constant_field():
    """tests/test_nonlinear_funs.py:126-152: c0 + c1 u + c2 u^2 on u = c."""
    n = 32
    f = nl.PolynomialNonlinearFun(1, n, dealiasing_fraction=2 / 3, coefficients=(0.5, -1.0, 2.0))
    u = np.full((1, n), 1.5)
    r = sp.ifft(f(sp.fft(u, num_spatial_dims=1)), num_spatial_dims=1, num_points=n)
    np.testing.assert_allclose(r, 0.5 - 1.5 + 2.0 * 2.25, atol=1e-12)


def test_dealiasing_zeroes_high_modes():
    """tests/test_nonlinear_funs.py:155-220: post-dealiased output."""
    n = 64
    d = sp.build_derivative_operator(1, 1.0, n)
    f = nl.ConvectionNonlinearFun(1, n, derivative_operator=d, dealiasing_fraction=2 / 3)
    uh = sp.fft(np.random.default_rng(1).standard_normal((1, n)), num_spatial_dims=1)
    out = f(uh)
    assert np.all(out[0, 21:] == 0)
    with pytest.raises(ValueError):
        nl.ZeroNonlinearFun(1, n).dealias(uh)


def test_multichannel_channel_check():
    """tests/test_nonlinear_funs.py:228-294: C must equal D."""
    d = sp.build_derivative_operator(2, 1.0, 16)
    f = nl.ConvectionNonlinearFun(2, 16, derivative_operator=d)
    with pytest.raises(ValueError):
        f(np.zeros((1, 16, 9), complex))


def test_kolmogorov_injection_value():
    """_vorticity_convection.py:166-176: -k_f gamma N^2/2 at index [0, k_f]."""
    n = 32
    d = sp.build_derivative_operator(2, 2 * np.pi, n)
    f = nl.VorticityConvection2dKolmogorov(2, n, derivative_operator=d, dealiasing_fraction=2 / 3,
                                           injection_mode=4, injection_scale=1.5)
    inj = f.injection[0]
    assert inj[0, 4] == pytest.approx(-4 * 1.5 * n * n / 2)
    assert np.count_nonzero(inj) == 1
    # in physical space the injection is -k_f gamma cos(k_f y)
    phys = sp.ifft(f.injection, num_spatial_dims=2, num_points=n)[0]
    y = sp.make_grid(2, 2 * np.pi, n)[1]
    np.testing.assert_allclose(phys, -4 * 1.5 * np.cos(4 * y), atol=1e-10)


def test_vorticity_convection_closed_form():
    """-(u . grad) omega for the Taylor-Green-like vortex omega = 2 sin x sin y
    is zero (u is tangent to the level sets); a sheared flow is not."""
    n = 32
    d = sp.build_derivative_operator(2, 2 * np.pi, n)
    f = nl.VorticityConvection2d(2, n, derivative_operator=d, dealiasing_fraction=2 / 3)
    g = sp.make_grid(2, 2 * np.pi, n)
    w = (2 * np.sin(g[0]) * np.sin(g[1]))[None]
    np.testing.assert_allclose(f(sp.fft(w, num_spatial_dims=2)), 0, atol=1e-9)
    w2 = (np.sin(g[0]) + np.cos(2 * g[1]))[None]
    assert np.abs(f(sp.fft(w2, num_spatial_dims=2))).max() > 1.0


def test_roots_of_unity():
    """tests/test_etdrk.py:8-17."""
    r = et.roots_of_unity(16)
    np.testing.assert_allclose(np.abs(r), 1, atol=1e-14)
    assert abs(r.sum()) < 1e-13


def _phi(z, k):
    """phi_k in 60-digit arithmetic (the closed forms cancel for small z)."""
    mpmath.mp.dps = 60
    z = mpmath.mpf(z)
    if k == 1:
        return (mpmath.e**z - 1) / z
    if k == 2:
        return (mpmath.e**z - 1 - z) / z**2
    if k == 3:
        return (mpmath.e**z - 1 - z - z**2 / 2) / z**3
    raise ValueError


@pytest.mark.parametrize("ldt", [-418.0, -30.0, -2.5, -0.3, -1e-3, 0.7, 5.0])
def test_contour_coefficients_match_phi_closed_forms(ldt):
    """The FP64 contour means equal dt * phi_k(L dt) (Kassam-Trefethen) and
    the ETDRK3/4 combinations (Cox-Matthews) for real L."""
    dt = 0.1
    L = np.array([[ldt / dt]])
    e2 = et.ETDRK2(dt, L, nl.ZeroNonlinearFun(1, 4))
    z = ldt
    rel = 1e-9
    assert e2._coef_1[0, 0] == pytest.approx(float(dt * _phi(z, 1)), rel=rel, abs=1e-14)
    assert e2._coef_2[0, 0] == pytest.approx(float(dt * _phi(z, 2)), rel=rel, abs=1e-14)
    e4 = et.ETDRK4(dt, L, nl.ZeroNonlinearFun(1, 4))
    p1, p2, p3 = (float(_phi(z, k)) for k in (1, 2, 3))
    p1h = float(_phi(z / 2, 1))
    assert e4._coef_1[0, 0] == pytest.approx(dt * 0.5 * p1h, rel=rel, abs=1e-14)
    # ETDRK4 (etdrk/_etdrk_4.py:175-196): c4 = phi1 - 3 phi2 + 4 phi3, c5 = phi2 - 2 phi3,
    # c6 = -phi2 + 4 phi3
    assert e4._coef_4[0, 0] == pytest.approx(dt * (p1 - 3 * p2 + 4 * p3), rel=1e-7, abs=1e-12)
    assert e4._coef_5[0, 0] == pytest.approx(dt * (p2 - 2 * p3), rel=1e-7, abs=1e-12)
    assert e4._coef_6[0, 0] == pytest.approx(dt * (-p2 + 4 * p3), rel=1e-7, abs=1e-12)
    e3 = et.ETDRK3(dt, L, nl.ZeroNonlinearFun(1, 4))
    assert e3._coef_3[0, 0] == pytest.approx(dt * (p1 - 3 * p2 + 4 * p3), rel=1e-7, abs=1e-12)
    assert e3._coef_4[0, 0] == pytest.approx(dt * 4 * (p2 - 2 * p3), rel=1e-7, abs=1e-12)
    assert e3._coef_5[0, 0] == pytest.approx(dt * (-p2 + 4 * p3), rel=1e-7, abs=1e-12)


def test_etdrk0_exact_diffusion():
    """tests/test_etdrk.py:20-45."""
    L, n, dt, nu = 10.0, 100, 0.1, 0.1

    class Diffusion(st.BaseStepper):
        def __init__(self):
            super().__init__(1, L, n, dt, num_channels=1, order=0)

        def _build_linear_operator(self, d):
            return nu * sp.build_laplace_operator(d)

        def _build_nonlinear_fun(self, d):
            return nl.ZeroNonlinearFun(1, n)

    k = 4 * 2 * np.pi / L
    g = sp.make_grid(1, L, n)
    u1 = Diffusion()(np.sin(k * g))
    np.testing.assert_allclose(u1, np.exp(-k * k * nu * dt) * np.sin(k * g), atol=1e-5)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_etdrk_convergence_order(order):
    """tests/test_etdrk.py:48-106: Burgers L=3, N=64, nu=0.02, sin IC; rate
    log2(e1/e2) > order - 0.25 against an ETDRK4 reference with 256 steps."""
    L, n, nu, dt0 = 3.0, 64, 0.02, 0.2
    g = sp.make_grid(1, L, n)
    u0 = np.sin(2 * np.pi * g / L)
    ref = st.repeat(st.Burgers(1, L, n, dt0 / 256, diffusivity=nu, order=4), 256)(u0)
    errs = []
    for r in (1, 2):
        s = st.Burgers(1, L, n, dt0 / r, diffusivity=nu, order=order)
        errs.append(np.sqrt(np.mean((st.repeat(s, r)(u0) - ref) ** 2)))
    assert np.log2(errs[0] / errs[1]) > order - 0.25


def test_orders_agree_on_burgers():
    """tests/test_builtin_solvers.py:929-953."""
    from oracle import ic

    u0 = ic.tfs(1, 2 * np.pi, 64, 5, np.random.default_rng(0))
    res = {o: st.Burgers(1, 2 * np.pi, 64, 0.01, diffusivity=0.1, order=o)(u0) for o in (1, 2, 3, 4)}
    for o in (3, 4):
        np.testing.assert_allclose(res[o], res[2], atol=5e-4)


def test_fp32_contour_close_to_fp64():
    """Forming the coefficients in float32 (the reference's default precision)
    differs from the float64 formation by ~1e-7 relative."""
    L = st.KuramotoSivashinskyConservative(1, 100.0, 256, 0.1).linear_operator
    a = et.ETDRK2(0.1, L, nl.ZeroNonlinearFun(1, 256), coef_dtype=np.float64)
    b = et.ETDRK2(0.1, L, nl.ZeroNonlinearFun(1, 256), coef_dtype=np.float32)
    np.testing.assert_allclose(b._coef_2, a._coef_2, rtol=5e-5, atol=1e-9)
