"""3-D on the GPU: rfftn / irfftn over the last three axes (_spectral.py:
614-721; z rows r2c, then y and x line FFTs) against numpy, and the linear
steppers in 3-D (ETDRK0, real and complex exp(L dt)) against the FP32 oracle
at steps 1 and 10 through single steps and the rollout, plus exact
transport."""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_fft3d_against_numpy(device, N):
    import torch

    print(f"[case] fft3d N={N}", flush=True)
    rng = np.random.default_rng(N)
    u = rng.standard_normal((2, 1, N, N, N)).astype(np.float32)
    uh = ex.spectral.fft(torch.tensor(u, device=device), num_spatial_dims=3)
    want = np.fft.rfftn(u.astype(np.float64), axes=(-3, -2, -1))
    err = np.abs(uh.cpu().numpy() - want).max() / np.abs(want).max()
    back = ex.spectral.ifft(uh, num_spatial_dims=3).cpu().numpy()
    shp = (2, N, N, N // 2 + 1)
    z = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(np.complex64)
    got_i = ex.spectral.ifft(torch.tensor(z, device=device), num_spatial_dims=3).cpu().numpy()
    want_i = np.fft.irfftn(z.astype(np.complex128), s=(N, N, N), axes=(-3, -2, -1))
    err_i = np.abs(got_i - want_i).max() / np.abs(want_i).max()
    print(f"  rfftn err {err:.2e}, roundtrip {np.abs(back - u).max():.2e}, irfftn err {err_i:.2e}", flush=True)
    assert err <= 2e-6 and err_i <= 2e-6
    np.testing.assert_allclose(back, u, atol=2e-5)


def _pair(name, N):
    L = 2.0
    if name == "advection":
        kw = dict(velocity=np.array([0.8, -0.3, 0.5]))
        return ex.stepper.Advection(3, L, N, 0.01, **kw), ost.Advection(3, L, N, 0.01, dtype=np.float32, **kw)
    if name == "advection_diffusion":
        kw = dict(velocity=-0.5, diffusivity=0.01)
        return (ex.stepper.AdvectionDiffusion(3, L, N, 0.01, **kw),
                ost.AdvectionDiffusion(3, L, N, 0.01, dtype=np.float32, **kw))
    if name == "dispersion":
        return (ex.stepper.Dispersion(3, L, N, 1e-4, dispersivity=0.2),
                ost.Dispersion(3, L, N, 1e-4, dispersivity=0.2, dtype=np.float32))
    if name == "diffusion":
        return (ex.stepper.Diffusion(3, L, N, 0.01, diffusivity=0.02),
                ost.Diffusion(3, L, N, 0.01, diffusivity=0.02, dtype=np.float32))
    return (ex.stepper.HyperDiffusion(3, L, N, 0.01, hyper_diffusivity=1e-4),
            ost.HyperDiffusion(3, L, N, 0.01, hyper_diffusivity=1e-4, dtype=np.float32))


@pytest.mark.parametrize("N", [16, 32])
@pytest.mark.parametrize("name", ["advection", "advection_diffusion", "dispersion", "diffusion", "hyper_diffusion"])
def test_linear_3d_parity(device, name, N):
    import torch

    print(f"[case] 3-D {name} N={N}", flush=True)
    gpu, cpu = _pair(name, N)
    u0, _ = ics(2, 3, 2.0, N, 3, seed=101)
    x = torch.tensor(u0, device=device)
    traj = gpu.rollout_batch(x, 10).cpu().numpy()
    ug, uc = x, u0
    for s in range(1, 11):
        ug = gpu.step(ug)
        uc = cpu.step_batch(uc)
        if s in (1, 10):
            e_step = max_rel_l2(ug.cpu().numpy(), uc)
            e_roll = max_rel_l2(traj[:, s - 1], uc)
            print(f"  step {s}: steps {e_step:.2e}, rollout {e_roll:.2e}", flush=True)
            assert e_step <= RTOL_STEP and e_roll <= RTOL_STEP


def test_advection_3d_exact_transport(device):
    import torch

    print("[case] 3-D exact transport", flush=True)
    L, N, dt, c = 2.0, 32, 0.05, np.array([0.7, -0.4, 0.3])
    x = np.linspace(0.0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = lambda x, y, z: np.sin(2 * np.pi * (2 * x + y - 3 * z) / L)  # noqa: E731
    st = ex.stepper.Advection(3, L, N, dt, velocity=c)
    u = st.rollout_batch(torch.tensor(f(X, Y, Z)[None, None], dtype=torch.float32, device=device), 20)
    got = u[0, -1, 0].cpu().numpy()
    t = 20 * dt
    np.testing.assert_allclose(got, f(X - c[0] * t, Y - c[1] * t, Z - c[2] * t), atol=2e-5)


def _check(gpu, cpu, u0, device, label):
    import torch

    x = torch.tensor(u0, device=device)
    traj = gpu.rollout_batch(x, 10).cpu().numpy()
    ug, uc = x, u0
    for s in range(1, 11):
        ug = gpu.step(ug)
        uc = cpu.step_batch(uc)
        if s in (1, 10):
            e_step = max_rel_l2(ug.cpu().numpy(), uc)
            e_roll = max_rel_l2(traj[:, s - 1], uc)
            print(f"  {label} step {s}: steps {e_step:.2e}, rollout {e_roll:.2e}", flush=True)
            assert e_step <= RTOL_STEP and e_roll <= RTOL_STEP


@pytest.mark.parametrize("N", [16, 32, 64])
@pytest.mark.parametrize("conservative,order", [(False, 2), (True, 2), (False, 1), (True, 4), (False, 3)])
def test_burgers_3d_parity(device, conservative, order, N):
    print(f"[case] 3-D Burgers conservative={conservative} order {order} N={N}", flush=True)
    kw = dict(diffusivity=0.01, conservative=conservative, order=order)
    gpu = ex.stepper.Burgers(3, 1.0, N, 1e-3, **kw)
    cpu = ost.Burgers(3, 1.0, N, 1e-3, dtype=np.float32, **kw)
    u0, _ = ics(2, 3, 1.0, N, 3, seed=111, channels=3)
    _check(gpu, cpu, u0, device, "burgers3d")


def _nl3d_pair(name, N):
    if name == "ks":
        return ex.stepper.KuramotoSivashinsky(3, 30.0, N, 0.05), ost.KuramotoSivashinsky(
            3, 30.0, N, 0.05, dtype=np.float32), 30.0, 1
    if name == "allen_cahn":
        return ex.stepper.reaction.AllenCahn(3, 1.0, N, 1e-3), ost.AllenCahn(3, 1.0, N, 1e-3,
                                                                             dtype=np.float32), 1.0, 1
    if name == "cahn_hilliard":
        return ex.stepper.reaction.CahnHilliard(3, 2.0, N, 1e-3), ost.CahnHilliard(3, 2.0, N, 1e-3,
                                                                                   dtype=np.float32), 2.0, 1
    if name == "gray_scott":
        kw = dict(diffusivity_1=2e-3, diffusivity_2=1e-3)
        return (ex.stepper.reaction.GrayScott(3, 1.0, N, 0.5, **kw),
                ost.GrayScott(3, 1.0, N, 0.5, dtype=np.float32, **kw), 1.0, 2)
    if name == "general":
        kw = dict(linear_coefficients=(0.0, 0.0, 0.01))
        return (ex.stepper.generic.GeneralNonlinearStepper(3, 2.0, N, 1e-3, nonlinear_coefficients=(0.3, -1.0, -0.5),
                                                           **kw),
                ost.GeneralStepper(3, 2.0, N, 1e-3, nonlinear_scales=(0.3, -1.0, -0.5), dtype=np.float32, **kw),
                2.0, 1)
    kw = dict(diffusivity=0.01, single_channel=True, conservative=name == "burgers_single")
    return (ex.stepper.Burgers(3, 1.0, N, 1e-3, **kw), ost.Burgers(3, 1.0, N, 1e-3, dtype=np.float32, **kw), 1.0, 1)


@pytest.mark.parametrize("name", ["ks", "allen_cahn", "cahn_hilliard", "gray_scott", "general", "burgers_single",
                                  "burgers_single_nc"])
def test_nonlinear_3d_parity(device, name):
    N = 32
    print(f"[case] 3-D {name} N={N}", flush=True)
    gpu, cpu, L, C = _nl3d_pair(name, N)
    u0, _ = ics(2, 3, L, N, 3, seed=112, channels=C)
    if name == "gray_scott":
        u0 = (np.array([0.7, 0.2], dtype=np.float32).reshape(1, 2, 1, 1, 1) + 0.2 * u0 / np.abs(u0).max()).astype(
            np.float32)
    _check(gpu, cpu, u0, device, name)


@pytest.mark.parametrize("name", ["ks", "cahn_hilliard", "gray_scott", "general"])
def test_nonlinear_3d_parity_n64(device, name):
    """N = 64, where the kernel that fuses the forward x lines with the
    combine and the ETDRK stage holds 8 points per lane (16 below)."""
    N = 64
    print(f"[case] 3-D {name} N={N}", flush=True)
    gpu, cpu, L, C = _nl3d_pair(name, N)
    u0, _ = ics(2, 3, L, N, 3, seed=113, channels=C)
    if name == "gray_scott":
        u0 = (np.array([0.7, 0.2], dtype=np.float32).reshape(1, 2, 1, 1, 1) + 0.2 * u0 / np.abs(u0).max()).astype(
            np.float32)
    _check(gpu, cpu, u0, device, name)


@pytest.mark.parametrize("N,order", [(16, 2), (32, 2), (32, 1), (32, 4), (64, 2), (64, 3)])
def test_navier_stokes_velocity_parity(device, N, order):
    """stepper/_navier_stokes.py:330-598: P(u x omega) (Leray-projected
    rotational convection) and, for N = 32 order 2, the Kolmogorov forcing."""
    print(f"[case] 3-D Navier-Stokes velocity N={N} order {order}", flush=True)
    L = 2 * np.pi
    kw = dict(diffusivity=0.02, order=order)
    gpu = ex.stepper.NavierStokesVelocity(3, L, N, 0.01, **kw)
    cpu = ost.NavierStokesVelocity(3, L, N, 0.01, dtype=np.float32, **kw)
    u0, _ = ics(2, 3, L, N, 3, seed=121, channels=3)
    _check(gpu, cpu, u0, device, "ns_velocity")
    if N == 32 and order == 2:
        kw.update(injection_mode=2, injection_scale=0.5, drag=-0.1)
        gpu = ex.stepper.KolmogorovFlowVelocity(3, L, N, 0.01, **kw)
        cpu = ost.NavierStokesVelocity(3, L, N, 0.01, dtype=np.float32, **kw)
        _check(gpu, cpu, u0, device, "kolmogorov_velocity")


def test_abc_flow_decays_viscously(device):
    import torch

    print("[case] 3-D ABC flow", flush=True)
    N, dt, nu, n = 32, 0.05, 0.02, 20
    x = np.linspace(0.0, 2 * np.pi, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u0 = np.stack([np.sin(Z) + 0.4 * np.cos(Y), 0.7 * np.sin(X) + np.cos(Z), 0.4 * np.sin(Y) + 0.7 * np.cos(X)])
    st = ex.stepper.NavierStokesVelocity(3, 2 * np.pi, N, dt, diffusivity=nu)
    got = st.repeat_batch(torch.tensor(u0[None], dtype=torch.float32, device=device), n)[0].cpu().numpy()
    np.testing.assert_allclose(got, np.exp(-nu * n * dt) * u0, atol=2e-5)


def test_step_fourier_nonlinear_and_repeated_3d(device):
    """3-D step_fourier (_base_stepper.py:222-239), the nonlinear function on
    spectra and RepeatedStepper (transform once, sub-step in Fourier space)
    against the oracle."""
    import torch

    from oracle import spectral as osp

    print("[case] 3-D step_fourier / nonlinear / repeated", flush=True)
    N, L = 16, 1.0
    kw = dict(diffusivity=0.01, order=2)
    gpu = ex.stepper.Burgers(3, L, N, 1e-3, **kw)
    cpu = ost.Burgers(3, L, N, 1e-3, dtype=np.float32, **kw)
    u0, _ = ics(2, 3, L, N, 3, seed=131, channels=3)
    uh = osp.fft(u0, num_spatial_dims=3).astype(np.complex64)
    got = gpu.step_fourier(torch.tensor(uh, device=device)).cpu().numpy()
    want = cpu.step_fourier(uh)
    e_sf = np.abs(got - want).max() / np.abs(want).max()
    got_n = gpu.nonlinear_fun(torch.tensor(uh, device=device)).cpu().numpy()
    want_n = cpu.nonlinear_fun(uh)
    e_nl = np.abs(got_n - want_n).max() / np.abs(want_n).max()
    rep = ex.RepeatedStepper(gpu, 5).step(torch.tensor(u0, device=device)).cpu().numpy()
    vh = uh
    for _ in range(5):
        vh = cpu.step_fourier(vh)
    want_r = osp.ifft(vh, num_spatial_dims=3, num_points=N)
    e_rep = max_rel_l2(rep, want_r)
    print(f"  step_fourier {e_sf:.2e}, nonlinear {e_nl:.2e}, repeated {e_rep:.2e}", flush=True)
    assert e_sf <= 1e-5 and e_nl <= 1e-5 and e_rep <= RTOL_STEP


def test_ensembles_3d_and_coupled_2d(device):
    """Per-trajectory linear operators with several channels: 3-D Burgers
    (three channels, shared L) and 2-D Gray-Scott (per-channel L) ensembles
    against the oracle's members."""
    import torch

    print("[case] 3-D Burgers / 2-D Gray-Scott ensembles", flush=True)
    nus = [0.005, 0.02]
    gpu = ex.ensemble(lambda nu: ex.stepper.Burgers(3, 1.0, 16, 1e-3, diffusivity=nu), nus)
    cpu = ost.Ensemble([ost.Burgers(3, 1.0, 16, 1e-3, diffusivity=nu, dtype=np.float32) for nu in nus],
                       dtype=np.float32)
    u0, _ = ics(2, 3, 1.0, 16, 3, seed=181, channels=3)
    _check(gpu, cpu, u0, device, "burgers3d_ensemble")
    d1 = [1e-3, 4e-3]
    gpu = ex.ensemble(lambda d: ex.stepper.reaction.GrayScott(2, 1.0, 64, 0.5, diffusivity_1=d, diffusivity_2=d / 2),
                      d1)
    cpu = ost.Ensemble([ost.GrayScott(2, 1.0, 64, 0.5, diffusivity_1=d, diffusivity_2=d / 2, dtype=np.float32)
                        for d in d1], dtype=np.float32)
    u0, _ = ics(2, 2, 1.0, 64, 4, seed=182, channels=2)
    u0 = (np.array([0.7, 0.2], dtype=np.float32).reshape(1, 2, 1, 1) + 0.2 * u0 / np.abs(u0).max()).astype(np.float32)
    _check(gpu, cpu, u0, device, "gray_scott_ensemble")
