"""Pins of the oracle's stepper layer (oracle/stepper.py) on the CPU:
rollout/repeat semantics (reference tests/test_utils.py), shape errors,
equivalences, the ensemble (vmap over constructors), and regression against
the committed fixtures."""
import os

import numpy as np
import pytest

from oracle import ic
from oracle import nonlin as nl
from oracle import spectral as sp
from oracle import stepper as st

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_rollout_equals_loop_and_repeat_equals_last():
    """tests/test_utils.py:25-81."""
    s = st.KuramotoSivashinskyConservative(1, 60.0, 64, 0.1)
    u0 = ic.tfs(1, 60.0, 64, 5, np.random.default_rng(0))
    tr = st.rollout(s, 5, include_init=True)(u0)
    assert tr.shape == (6, 1, 64)
    u = u0
    for i in range(5):
        u = s(u)
        np.testing.assert_allclose(tr[i + 1], u, atol=1e-12)
    np.testing.assert_allclose(st.repeat(s, 5)(u0), tr[-1], atol=1e-12)
    assert st.rollout(s, 5)(u0).shape == (5, 1, 64)


def test_call_shape_error():
    """_base_stepper.py:263-270."""
    s = st.Burgers(1, 1.0, 32, 0.01)
    with pytest.raises(ValueError, match="Expected shape"):
        s(np.zeros((2, 32)))


def test_linear_operator_shape_error_and_order():
    """_base_stepper.py:111-116 and :156-157."""

    class Bad(st.BaseStepper):
        def _build_linear_operator(self, d):
            return np.zeros((3, 17))

        def _build_nonlinear_fun(self, d):
            return nl.ZeroNonlinearFun(1, 32)

    with pytest.raises(ValueError, match="linear operator"):
        Bad(1, 1.0, 32, 0.1, num_channels=1, order=0)
    with pytest.raises(NotImplementedError):
        st.Burgers(1, 1.0, 32, 0.1, order=5)


def test_ks_conservative_equals_general_stepper():
    """tests/test_builtin_solvers.py:126-222: KS-conservative ==
    General(linear (0, 0, -1, 0, -1), nonlinear (0, -1, 0))."""
    u0 = ic.tfs(1, 30.0, 64, 4, np.random.default_rng(2))
    a = st.KuramotoSivashinskyConservative(1, 30.0, 64, 0.05)(u0)
    b = st.GeneralStepper(1, 30.0, 64, 0.05, linear_coefficients=(0.0, 0.0, -1.0, 0.0, -1.0),
                          nonlinear_scales=(0.0, -1.0, 0.0))(u0)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)


def test_ks_equals_general_gradient_norm():
    """tests/test_builtin_solvers.py:225-299."""
    u0 = ic.tfs(1, 30.0, 64, 4, np.random.default_rng(3))
    a = st.KuramotoSivashinsky(1, 30.0, 64, 0.05)(u0)
    b = st.GeneralStepper(1, 30.0, 64, 0.05, linear_coefficients=(0.0, 0.0, -1.0, 0.0, -1.0),
                          nonlinear_scales=(0.0, 0.0, -1.0))(u0)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)


def test_navier_stokes_kolmogorov_without_forcing():
    """KolmogorovFlowVorticity with zero injection equals NavierStokesVorticity
    with the same drag (tests/test_builtin_solvers.py:878-895 style)."""
    u0 = ic.tfs(2, 2 * np.pi, 32, 3, np.random.default_rng(4))
    a = st.KolmogorovFlowVorticity(2, 2 * np.pi, 32, 0.01, diffusivity=0.01, injection_scale=0.0)(u0)
    b = st.NavierStokesVorticity(2, 2 * np.pi, 32, 0.01, diffusivity=0.01, drag=-0.1)(u0)
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_burgers_energy_decays():
    """tests/test_builtin_solvers.py:424-489: viscous Burgers loses energy."""
    u0 = ic.tfs(1, 1.0, 128, 4, np.random.default_rng(5))
    s = st.Burgers(1, 1.0, 128, 1e-3, diffusivity=1e-2)
    e = [np.sum(u0**2)]
    u = u0
    for _ in range(20):
        u = s(u)
        e.append(np.sum(u**2))
    assert all(b < a for a, b in zip(e, e[1:]))


def test_ensemble_equals_members():
    """The vmap over constructors (per-trajectory viscosity) equals stepping
    each member on its own."""
    nus = [1e-3, 5e-3, 2e-2]
    members = [st.Burgers(1, 1.0, 64, 1e-3, diffusivity=n) for n in nus]
    ens = st.Ensemble(members)
    u0, _ = ic.tfs_batch(3, 1, 1.0, 64, 4, 7)
    out = ens.step_batch(u0)
    for i, m in enumerate(members):
        np.testing.assert_allclose(out[i], m(u0[i]), atol=1e-13)


def test_batch_step_equals_single():
    """step_batch is the vmap of __call__ for every nonlinear kind."""
    u0, _ = ic.tfs_batch(3, 1, 20.0, 64, 4, 8)
    for s in (st.KuramotoSivashinskyConservative(1, 20.0, 64, 0.05), st.KuramotoSivashinsky(1, 20.0, 64, 0.05),
              st.Burgers(1, 20.0, 64, 0.05, diffusivity=0.5)):
        b = s.step_batch(u0)
        for i in range(3):
            np.testing.assert_allclose(b[i], s(u0[i]), atol=1e-13)


def test_nrmse():
    """metrics/_spatial.py:537-586: sum over channels of relative L2 errors."""
    a = np.ones((2, 8))
    b = np.ones((2, 8)) * 2
    assert st.nrmse(a, b) == pytest.approx(2 * 0.5)


@pytest.mark.parametrize("name,cls,args,nu", [
    ("ks_cons_1d", "KuramotoSivashinskyConservative", (1, 100.0, 256, 0.1), False),
    ("ks_1d", "KuramotoSivashinsky", (1, 60.0, 128, 0.05), False),
    ("burgers_1d", "Burgers", (1, 1.0, 1024, 1e-3), True),
])
def test_oracle_regression_fixtures(name, cls, args, nu):
    """The oracle reproduces the committed fixtures (written by
    tests/golden/make_golden.py) exactly in FP64 and to FP32 rounding."""
    f = np.load(os.path.join(GOLDEN, f"oracle_{name}.npz"))
    C = getattr(st, cls)
    for tag, dtype in (("f64", np.float64), ("f32", np.float32)):
        if nu:
            s = st.Ensemble([C(*args, diffusivity=float(n), dtype=dtype) for n in f["nu"]], dtype=dtype)
        else:
            s = C(*args, dtype=dtype)
        u = f["u0"].astype(dtype)
        for k in range(1, 11):
            u = s.step_batch(u)
            if k in (1, 10):
                np.testing.assert_allclose(u, f[f"{tag}_step{k}"], rtol=0, atol=1e-12 if tag == "f64" else 1e-5)
    # FP32 vs FP64 at 10 steps: the FP32 rounding floor the GPU path is held to
    for b in range(f["u0"].shape[0]):
        assert st.nrmse(f["f32_step10"][b], f["f64_step10"][b]) < 1e-5


def test_ic_recipe():
    """Zero mean, unit max-abs, seeded and reproducible."""
    u, _ = ic.tfs_batch(2, 1, 1.0, 256, 5, 1)
    v, _ = ic.tfs_batch(2, 1, 1.0, 256, 5, 1)
    np.testing.assert_array_equal(u, v)
    assert np.allclose(u.mean(axis=-1), 0, atol=1e-12)
    assert np.allclose(np.abs(u).max(axis=-1), 1)
    u2 = ic.tfs(2, 1.0, 32, 3, np.random.default_rng(0))
    assert u2.shape == (1, 32, 32) and abs(u2.mean()) < 1e-12
    # band-limited to the cutoff
    uh = sp.fft(u2, num_spatial_dims=2)
    wn = sp.build_wavenumbers(2, 32)
    assert np.abs(uh[0][np.max(np.abs(wn), axis=0) > 3]).max() < 1e-9


def test_rollout_resume_bit_exact():
    """Checkpoint / resume: 100 steps, then 100 more from the (step, state)
    snapshot, equal a continuous 200-step rollout bit for bit."""
    s = st.KuramotoSivashinskyConservative(1, 100.0, 64, 0.1, dtype=np.float32)
    u0 = ic.tfs(1, 100.0, 64, 5, np.random.default_rng(12)).astype(np.float32)
    full = st.rollout(s, 200)(u0)
    first = st.rollout(s, 100)(u0)
    snap = (100, first[-1].copy())
    second = st.rollout(s, 100)(snap[1])
    np.testing.assert_array_equal(np.concatenate([first, second]), full)


# ---------------------------------------------------------------- linear steppers and KdV
# Analytic pins of the linear operators (ETDRK0 is the exact exponential) and
# of the KdV signs (a travelling soliton).
def _grid(L, N):
    return np.linspace(0.0, L, N, endpoint=False)


@pytest.mark.parametrize("velocity", [0.7, -1.3])
def test_advection_exact_shift(velocity):
    L, N, dt = 2.0, 64, 0.1
    x = _grid(L, N)
    adv = st.Advection(1, L, N, dt, velocity=velocity)
    u = adv.step(np.sin(2 * np.pi * 3 * x / L)[None])
    np.testing.assert_allclose(u[0], np.sin(2 * np.pi * 3 * (x - velocity * dt) / L), atol=1e-12)


def test_diffusion_and_hyperdiffusion_exact_decay():
    L, N, dt, k = 3.0, 48, 0.2, 4
    x = _grid(L, N)
    kap = 2 * np.pi * k / L
    u0 = np.cos(kap * x)[None]
    np.testing.assert_allclose(st.Diffusion(1, L, N, dt, diffusivity=0.05).step(u0)[0],
                               np.exp(-0.05 * kap**2 * dt) * u0[0], atol=1e-12)
    np.testing.assert_allclose(st.HyperDiffusion(1, L, N, dt, hyper_diffusivity=1e-3).step(u0)[0],
                               np.exp(-1e-3 * kap**4 * dt) * u0[0], atol=1e-12)
    ad = st.AdvectionDiffusion(1, L, N, dt, velocity=0.5, diffusivity=0.05).step(u0)[0]
    np.testing.assert_allclose(ad, np.exp(-0.05 * kap**2 * dt) * np.cos(kap * (x - 0.5 * dt)), atol=1e-12)


def test_dispersion_exact_phase():
    # u_t = a u_xxx: e^{i k x} -> e^{i k (x - a k^2 t)}
    L, N, dt, a, k = 2 * np.pi, 64, 0.05, 0.3, 3
    x = _grid(L, N)
    u = st.Dispersion(1, L, N, dt, dispersivity=a).step(np.sin(k * x)[None])
    np.testing.assert_allclose(u[0], np.sin(k * (x - a * k**2 * dt)), atol=1e-12)


def test_kdv_soliton():
    """u_t - 6 u u_x + u_xxx = 0 (convection_scale -6, dispersivity 1, no
    hyper-diffusion) carries u = -(c/2) sech^2(sqrt(c)/2 (x - c t - x0))."""
    L, N, dt, c, x0, n = 40.0, 256, 1e-3, 1.0, 15.0, 400
    x = _grid(L, N)
    sol = lambda t: -0.5 * c / np.cosh(0.5 * np.sqrt(c) * (x - c * t - x0)) ** 2  # noqa: E731
    kdv = st.KortewegDeVries(1, L, N, dt, hyper_diffusivity=0.0, dealiasing_fraction=1.0)
    u = sol(0.0)[None]
    for _ in range(n):
        u = kdv.step(u)
    # 2.5e-4 with the reference's convection_scale -6; the opposite sign moves it the wrong way (0.28)
    assert np.abs(u[0] - sol(n * dt)).max() < 2e-3


# ---------------------------------------------------------------- polynomial reactions
# A spatially uniform state follows the reaction ODE u' = L(0) u + p(u).
def _uniform_run(stepper, u0, n):
    u = np.full((1, stepper.num_points), u0)
    for _ in range(n):
        u = stepper.step(u)
    assert np.ptp(u) < 1e-12
    return float(u[0, 0])


def test_fisher_kpp_logistic():
    r, u0, dt, n = 1.0, 0.1, 0.01, 100
    got = _uniform_run(st.FisherKPP(1, 1.0, 32, dt, reactivity=r), u0, n)
    t = n * dt
    exact = u0 * np.exp(r * t) / (1 - u0 + u0 * np.exp(r * t))
    assert abs(got - exact) < 2e-5


def test_allen_cahn_cubic():
    u0, dt, n = 0.2, 0.01, 100
    got = _uniform_run(st.AllenCahn(1, 1.0, 32, dt), u0, n)  # u' = u - u^3
    t = n * dt
    exact = u0 * np.exp(t) / np.sqrt(1 + u0**2 * (np.exp(2 * t) - 1))
    assert abs(got - exact) < 2e-5


def test_swift_hohenberg_uniform():
    from scipy.integrate import solve_ivp

    u0, dt, n = 0.5, 0.01, 100
    got = _uniform_run(st.SwiftHohenberg(1, 10.0, 32, dt), u0, n)  # u' = (0.7 - 1) u + u^2 - u^3
    ref = solve_ivp(lambda t, u: -0.3 * u + u**2 - u**3, (0.0, n * dt), [u0], rtol=1e-12, atol=1e-14).y[0, -1]
    assert abs(got - ref) < 2e-5


# ---- 2-D: the same closed forms on a 2-D grid (tests/test_validation.py:13-211
# covers advection and diffusion in 1/2/3-D)
def _grid2(L, N):
    x = _grid(L, N)
    return np.meshgrid(x, x, indexing="ij")


def test_advection_2d_exact_shift():
    L, N, dt, c = 2.0, 32, 0.1, (0.7, -0.4)
    X, Y = _grid2(L, N)
    f = lambda x, y: np.sin(2 * np.pi * (3 * x + 2 * y) / L)  # noqa: E731
    u = st.Advection(2, L, N, dt, velocity=np.array(c)).step(f(X, Y)[None])
    np.testing.assert_allclose(u[0], f(X - c[0] * dt, Y - c[1] * dt), atol=1e-12)


def test_diffusion_2d_exact_decay():
    L, N, dt, nu = 3.0, 32, 0.2, 0.05
    X, Y = _grid2(L, N)
    kx, ky = 2 * np.pi * 3 / L, 2 * np.pi * 2 / L
    u0 = np.cos(kx * X) * np.sin(ky * Y)
    u = st.Diffusion(2, L, N, dt, diffusivity=nu).step(u0[None])
    np.testing.assert_allclose(u[0], np.exp(-nu * (kx**2 + ky**2) * dt) * u0, atol=1e-12)


def test_allen_cahn_2d_uniform():
    u0, dt, n = 0.2, 0.01, 100
    ac = st.AllenCahn(2, 1.0, 16, dt)
    u = np.full((1, 16, 16), u0)
    for _ in range(n):
        u = ac.step(u)
    assert np.ptp(u) < 1e-12
    t = n * dt
    assert abs(u[0, 0, 0] - u0 * np.exp(t) / np.sqrt(1 + u0**2 * (np.exp(2 * t) - 1))) < 2e-5


@pytest.mark.parametrize("conservative", [False, True])
def test_burgers_2d_reduces_to_1d(conservative):
    """A velocity field (f(x), 0) or (0, g(y)) of 2-D Burgers evolves as 1-D
    Burgers along its axis (the y-independent modes have ky = 0 only)."""
    L, N, dt, nu = 1.0, 32, 0.01, 0.02
    x = _grid(L, N)
    f = 0.6 * np.sin(2 * np.pi * x / L) + 0.3 * np.cos(2 * np.pi * 3 * x / L)
    one = st.Burgers(1, L, N, dt, diffusivity=nu, conservative=conservative).step(f[None])[0]
    two = st.Burgers(2, L, N, dt, diffusivity=nu, conservative=conservative)
    ux = two.step(np.stack([np.repeat(f[:, None], N, axis=1), np.zeros((N, N))]))
    np.testing.assert_allclose(ux[0], np.repeat(one[:, None], N, axis=1), atol=1e-12)
    np.testing.assert_allclose(ux[1], 0.0, atol=1e-12)
    vy = two.step(np.stack([np.zeros((N, N)), np.repeat(f[None, :], N, axis=0)]))
    np.testing.assert_allclose(vy[1], np.repeat(one[None, :], N, axis=0), atol=1e-12)
    np.testing.assert_allclose(vy[0], 0.0, atol=1e-12)


@pytest.mark.parametrize("conservative", [False, True])
def test_burgers_2d_axis_swap_equivariance(conservative):
    """Swapping the axes together with the velocity components commutes with
    a 2-D Burgers step."""
    from oracle import ic as oic

    u0, _ = oic.tfs_batch(1, 2, 1.0, 32, 4, 5, num_channels=2)
    u0 = u0[0]
    bu = st.Burgers(2, 1.0, 32, 0.01, diffusivity=0.02, conservative=conservative)
    swap = lambda u: np.stack([u[1].T, u[0].T])  # noqa: E731
    np.testing.assert_allclose(bu.step(swap(u0)), swap(bu.step(u0)), atol=1e-12)


def test_general_and_single_channel_convection_2d_reduce_to_1d():
    L, N, dt = 2.0, 32, 0.01
    x = _grid(L, N)
    f = 0.5 * np.sin(2 * np.pi * x / L) + 0.2 * np.sin(2 * np.pi * 4 * x / L)
    kw = dict(linear_coefficients=(0.0, 0.0, 0.01), nonlinear_scales=(0.3, -1.0, -0.5))
    one = st.GeneralStepper(1, L, N, dt, **kw).step(f[None])[0]
    two = st.GeneralStepper(2, L, N, dt, **kw).step(np.repeat(f[:, None], N, axis=1)[None])[0]
    np.testing.assert_allclose(two, np.repeat(one[:, None], N, axis=1), atol=1e-12)
    kw = dict(diffusivity=0.01, single_channel=True, conservative=True)
    one = st.Burgers(1, L, N, dt, **kw).step(f[None])[0]
    two = st.Burgers(2, L, N, dt, **kw).step(np.repeat(f[None, :], N, axis=0)[None])[0]
    np.testing.assert_allclose(two, np.repeat(one[None, :], N, axis=0), atol=1e-12)


@pytest.mark.parametrize("D", [1, 2])
def test_cahn_hilliard_linearised_growth(D):
    """u = a + eps cos(kappa x): u^3 = a^3 + 3 a^2 eps cos + O(eps^2), so the
    cosine amplitude grows at -nu kappa^2 (3 c3 a^2 + c1 + gamma kappa^2)."""
    L, N, dt, n, nu, gam, c1, c3, a, eps, m = 2.0, 32, 1e-3, 200, 1e-2, 1e-3, -1.0, 1.0, 0.3, 1e-7, 2
    x = _grid(L, N)
    kap = 2 * np.pi * m / L
    prof = np.cos(kap * x)
    u = (a + eps * (prof if D == 1 else np.repeat(prof[:, None], N, axis=1)))[None]
    ch = st.CahnHilliard(D, L, N, dt, diffusivity=nu, gamma=gam, first_order_coefficient=c1,
                         third_order_coefficient=c3, order=2)
    for _ in range(n):
        u = ch.step(u)
    lam = -nu * kap**2 * (3 * c3 * a**2 + c1 + gam * kap**2)
    row = u[0] if D == 1 else u[0][:, 0]
    amp = 2 * np.abs(np.fft.rfft(row)[m]) / N
    assert abs(row.mean() - a) < 1e-12
    assert abs(amp / (eps * np.exp(lam * n * dt)) - 1) < 1e-4
    # the wrong sign of the nonlinear Laplacian moves the rate by 2 * 3 nu kappa^2 c3 a^2
    assert abs(np.exp(6 * nu * kap**2 * c3 * a**2 * n * dt) - 1) > 1e-3


def _uniform_channels(stepper, values, n):
    u = np.stack([np.full(stepper.num_points, v) for v in values])
    for _ in range(n):
        u = stepper.step(u)
    assert np.ptp(u, axis=1).max() < 1e-12
    return u[:, 0]


def test_gray_scott_uniform_ode():
    from scipy.integrate import solve_ivp

    f, k, dt, n, y0 = 0.04, 0.06, 0.01, 200, (0.6, 0.3)
    got = _uniform_channels(st.GrayScott(1, 1.0, 16, dt, feed_rate=f, kill_rate=k), y0, n)
    rhs = lambda t, y: [f * (1 - y[0]) - y[0] * y[1] ** 2, -(f + k) * y[1] + y[0] * y[1] ** 2]  # noqa: E731
    ref = solve_ivp(rhs, (0.0, n * dt), list(y0), rtol=1e-12, atol=1e-14).y[:, -1]
    np.testing.assert_allclose(got, ref, atol=1e-7)


def test_gray_scott_linearised_decay_per_channel():
    """Around (1, 0) a small cosine decays at nu1 kappa^2 + f in channel 0 and
    nu2 kappa^2 + f + k in channel 1 (per-channel diffusivities)."""
    L, N, dt, n, nu1, nu2, f, k, eps, m = 1.0, 32, 0.01, 100, 2e-3, 5e-4, 0.04, 0.06, 1e-7, 3
    x = _grid(L, N)
    kap = 2 * np.pi * m / L
    c = np.cos(kap * x)
    u = np.stack([1 + eps * c, eps * c])
    gs = st.GrayScott(1, L, N, dt, diffusivity_1=nu1, diffusivity_2=nu2, feed_rate=f, kill_rate=k)
    for _ in range(n):
        u = gs.step(u)
    t = n * dt
    amp = 2 * np.abs(np.fft.rfft(u, axis=1)[:, m]) / N
    np.testing.assert_allclose(amp / eps, [np.exp(-(nu1 * kap**2 + f) * t), np.exp(-(nu2 * kap**2 + f + k) * t)],
                               rtol=1e-5)


def test_belousov_zhabotinsky_uniform_ode():
    from scipy.integrate import solve_ivp

    dt, n, y0 = 0.005, 200, (0.5, 0.2, 0.3)
    got = _uniform_channels(st.BelousovZhabotinsky(1, 1.0, 16, dt), y0, n)
    rhs = lambda t, y: [y[0] + y[1] - y[0] * y[1] - y[0] ** 2, y[2] - y[1] - y[0] * y[1], y[0] - y[2]]  # noqa: E731
    ref = solve_ivp(rhs, (0.0, n * dt), list(y0), rtol=1e-12, atol=1e-14).y[:, -1]
    np.testing.assert_allclose(got, ref, atol=2e-6)


def test_linear_3d_exact():
    """3-D advection and diffusion against their closed forms
    (tests/test_validation.py:13-211 covers 1/2/3-D)."""
    L, N, dt, c, nu = 2.0, 16, 0.1, (0.7, -0.4, 0.3), 0.05
    x = _grid(L, N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = lambda x, y, z: np.sin(2 * np.pi * (2 * x + y - 3 * z) / L)  # noqa: E731
    u = st.Advection(3, L, N, dt, velocity=np.array(c)).step(f(X, Y, Z)[None])
    np.testing.assert_allclose(u[0], f(X - c[0] * dt, Y - c[1] * dt, Z - c[2] * dt), atol=1e-12)
    k2 = (2 * np.pi / L) ** 2 * (4 + 1 + 9)
    u = st.Diffusion(3, L, N, dt, diffusivity=nu).step(f(X, Y, Z)[None])
    np.testing.assert_allclose(u[0], np.exp(-nu * k2 * dt) * f(X, Y, Z), atol=1e-12)


@pytest.mark.parametrize("conservative", [False, True])
def test_burgers_3d_reduces_to_1d(conservative):
    L, N, dt, nu = 1.0, 16, 0.01, 0.02
    x = _grid(L, N)
    f = 0.6 * np.sin(2 * np.pi * x / L) + 0.3 * np.cos(2 * np.pi * 3 * x / L)
    one = st.Burgers(1, L, N, dt, diffusivity=nu, conservative=conservative).step(f[None])[0]
    three = st.Burgers(3, L, N, dt, diffusivity=nu, conservative=conservative)
    z = np.zeros((N, N, N))
    out = three.step(np.stack([z, z, np.broadcast_to(f, (N, N, N))]))  # w(z) along the last axis
    np.testing.assert_allclose(out[2], np.broadcast_to(one, (N, N, N)), atol=1e-12)
    np.testing.assert_allclose(out[:2], 0.0, atol=1e-12)


def test_ks_and_general_3d_reduce_to_1d():
    L, N, dt = 30.0, 16, 0.05
    x = _grid(L, N)
    f = 0.5 * np.sin(2 * np.pi * x / L) + 0.2 * np.sin(2 * np.pi * 3 * x / L)
    one = st.KuramotoSivashinsky(1, L, N, dt).step(f[None])[0]
    three = st.KuramotoSivashinsky(3, L, N, dt).step(np.broadcast_to(f[:, None, None], (N, N, N))[None])[0]
    np.testing.assert_allclose(three, np.broadcast_to(one[:, None, None], (N, N, N)), atol=1e-12)
    kw = dict(linear_coefficients=(0.0, 0.0, 0.01), nonlinear_scales=(0.3, -1.0, -0.5))
    one = st.GeneralStepper(1, 2.0, N, 0.01, **kw).step(f[None])[0]
    three = st.GeneralStepper(3, 2.0, N, 0.01, **kw).step(np.broadcast_to(f[None, :, None], (N, N, N))[None])[0]
    np.testing.assert_allclose(three, np.broadcast_to(one[None, :, None], (N, N, N)), atol=1e-12)


def _abc(N, A=1.0, B=0.7, C=0.4):
    x = _grid(2 * np.pi, N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z), C * np.sin(Y) + B * np.cos(X)])


def test_navier_stokes_velocity_abc_flow_decays_viscously():
    """The ABC flow is a Beltrami field (curl u = u), so u x omega = 0 and
    every mode (|k| = 1) decays as exp(-nu t)."""
    N, dt, nu, n = 16, 0.05, 0.02, 10
    u0 = _abc(N)
    ns = st.NavierStokesVelocity(3, 2 * np.pi, N, dt, diffusivity=nu)
    u = u0
    for _ in range(n):
        u = ns.step(u)
    np.testing.assert_allclose(u, np.exp(-nu * n * dt) * u0, atol=1e-12)


def test_leray_projection_properties():
    from oracle import nonlin as onl
    from oracle import spectral as osp

    N = 16
    d = osp.build_derivative_operator(3, 2 * np.pi, N)
    P = onl.Leray(d)
    rng = np.random.default_rng(0)
    u_hat = osp.fft(rng.standard_normal((3, N, N, N)), num_spatial_dims=3)
    pu = P(u_hat)
    np.testing.assert_allclose(P(pu), pu, atol=1e-9)  # idempotent
    assert np.abs(np.sum(d * pu, axis=0)).max() < 1e-9 * np.abs(u_hat).max()  # divergence-free
    phi_hat = osp.fft(rng.standard_normal((1, N, N, N)), num_spatial_dims=3)
    assert np.abs(P(d * phi_hat)).max() < 1e-9 * np.abs(d * phi_hat).max()  # gradients vanish


def test_navier_stokes_velocity_keeps_divergence_free():
    from oracle import nonlin as onl
    from oracle import spectral as osp

    N, L = 16, 2 * np.pi
    d = osp.build_derivative_operator(3, L, N)
    rng = np.random.default_rng(1)
    k = osp.build_wavenumbers(3, N)
    low = (np.abs(k) <= 3).all(axis=0)
    u_hat = onl.Leray(d)(osp.fft(rng.standard_normal((3, N, N, N)), num_spatial_dims=3) * low)
    u = osp.ifft(u_hat, num_spatial_dims=3, num_points=N)
    ns = st.NavierStokesVelocity(3, L, N, 0.01, diffusivity=0.01)
    for _ in range(5):
        u = ns.step(u)
    uh = osp.fft(u, num_spatial_dims=3)
    assert np.abs(np.sum(d * uh, axis=0)).max() < 1e-10 * np.abs(uh).max()


def test_kolmogorov_velocity_from_rest():
    """From rest, N(0) is the forcing and the second stage's u x omega is a
    gradient (u = (f(y), 0, 0)), so one ETDRK2 step gives c1 * forcing."""
    from oracle import spectral as osp

    N, L, dt, nu, kf = 16, 2 * np.pi, 0.1, 0.05, 2
    ks = st.NavierStokesVelocity(3, L, N, dt, diffusivity=nu, injection_mode=kf, injection_scale=0.5)
    u1 = ks.step(np.zeros((3, N, N, N)))
    f_hat = ks.nonlinear_fun.injection
    lam = -nu * (2 * np.pi / L * kf) ** 2
    # the state is real: compare after the same physical round trip
    want = osp.ifft((np.exp(lam * dt) - 1) / lam * f_hat, num_spatial_dims=3, num_points=N)
    np.testing.assert_allclose(u1, want, atol=1e-9 * np.abs(want).max())
    assert np.abs(want).max() > 0.01


@pytest.mark.parametrize("D", [1, 2])
def test_wave_exact_standing_wave_and_drift(D):
    """h = cos(kappa x), v = 0 gives h = cos(kappa x) cos(c kappa t), v = -c kappa
    cos(kappa x) sin(c kappa t); a uniform (a, b) drifts as h = a + b t."""
    L, N, dt, c, m, n = 2.0, 16, 0.07, 1.3, 3, 9
    x = _grid(L, N)
    kap = 2 * np.pi * m / L
    prof = np.cos(kap * x) if D == 1 else np.cos(kap * x)[:, None] * np.ones(N)
    wave = st.Wave(D, L, N, dt, speed_of_sound=c)
    u = np.stack([prof, np.zeros_like(prof)])
    for _ in range(n):
        u = wave.step(u)
    t = n * dt
    np.testing.assert_allclose(u[0], prof * np.cos(c * kap * t), atol=1e-12)
    np.testing.assert_allclose(u[1], -c * kap * prof * np.sin(c * kap * t), atol=1e-11)
    u = np.stack([np.full_like(prof, 0.4), np.full_like(prof, -0.3)])
    for _ in range(n):
        u = wave.step(u)
    np.testing.assert_allclose(u[0], 0.4 - 0.3 * t, atol=1e-12)
    np.testing.assert_allclose(u[1], -0.3, atol=1e-12)


def test_wave_diagonalisation_equals_product_form():
    """The reference's rotation (w +- v)/sqrt 2, w = i c |k| h, advanced by
    exp(+- i c |k| dt) and rotated back equals h' = cos h + sin/(c|k|) v,
    v' = -c|k| sin h + cos v (the GPU's form), and h' = h + dt v at k = 0."""
    from oracle import spectral as osp

    N, L, dt, c = 32, 2.0, 0.03, 1.3
    wave = st.Wave(2, L, N, dt, speed_of_sound=c)
    rng = np.random.default_rng(0)
    uh = osp.fft(rng.standard_normal((2, N, N)), num_spatial_dims=2)
    got = wave.step_fourier(uh)
    ck = c * wave.knorm[0]
    th = ck * dt
    with np.errstate(divide="ignore", invalid="ignore"):
        bh = np.where(ck == 0, dt, np.sin(th) / np.where(ck == 0, 1.0, ck))
    want = np.stack([np.cos(th) * uh[0] + bh * uh[1], -ck * np.sin(th) * uh[0] + np.cos(th) * uh[1]])
    np.testing.assert_allclose(got, want, atol=1e-10 * np.abs(uh).max())


@pytest.mark.parametrize("D", [1, 2, 3])
def test_poisson_closed_form(D):
    """-Laplace u = f for f = sin(2 pi (k . x) / L) gives u = f / |kappa|^2
    (tests/test_validation.py:214-292); order 4 divides by |kappa|^4 summed per
    axis; the mean is dropped."""
    L, N = 3.0, 16
    x = _grid(L, N)
    grids = np.meshgrid(*([x] * D), indexing="ij")
    ks = [2, -1, 3][:D]
    arg = sum(k * g for k, g in zip(ks, grids)) * 2 * np.pi / L
    f = np.sin(arg) + 0.7
    kap2 = sum((2 * np.pi * k / L) ** 2 for k in ks)
    u = st.Poisson(D, L, N).step(f[None])[0]
    np.testing.assert_allclose(u, np.sin(arg) / kap2, atol=1e-12)
    kap4 = sum((2 * np.pi * k / L) ** 4 for k in ks)
    u4 = st.Poisson(D, L, N, order=4).step(f[None])[0]
    np.testing.assert_allclose(u4, -np.sin(arg) / kap4, atol=1e-12)


def test_forced_diffusion_closed_form():
    """ForcedStepper around exact diffusion: u' = E (u + dt f) per mode."""
    L, N, dt, nu, m = 2.0, 32, 0.05, 0.03, 2
    x = _grid(L, N)
    kap = 2 * np.pi * m / L
    u, f = np.cos(kap * x)[None], 0.5 * np.cos(kap * x)[None]
    got = st.ForcedStepper(st.Diffusion(1, L, N, dt, diffusivity=nu)).step_batch(u, f)
    np.testing.assert_allclose(got, np.exp(-nu * kap**2 * dt) * (1 + 0.5 * dt) * u, atol=1e-12)
