"""Non-power-of-two 1-D resolutions N = 48 and N = 50 (the default of the
difficulty parametrisations, stepper/generic/_convection.py:248, and a size
the reference's own tests use): the single-lane radix-3 and radix-5
transforms against the FP32 oracle, through single steps and the multi-step
kernel, for real and complex linear operators and ETDRK orders 0-4; the
spectral transforms against numpy; the difficulty steppers at their default
resolution. N = 16 (a single-lane radix-8 transform) shares the register
path and is checked alongside."""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu

SIZES = [16, 48, 50]


def _pair(name, N, order):
    if name == "burgers":
        kw = dict(diffusivity=0.02, order=order)
        return ex.stepper.Burgers(1, 1.0, N, 1e-3, **kw), ost.Burgers(1, 1.0, N, 1e-3, dtype=np.float32, **kw), 1.0
    if name == "ks_cons":
        return (ex.stepper.KuramotoSivashinskyConservative(1, 30.0, N, 0.05, order=order),
                ost.KuramotoSivashinskyConservative(1, 30.0, N, 0.05, order=order, dtype=np.float32), 30.0)
    if name == "ks":
        return (ex.stepper.KuramotoSivashinsky(1, 30.0, N, 0.05, order=order),
                ost.KuramotoSivashinsky(1, 30.0, N, 0.05, order=order, dtype=np.float32), 30.0)
    if name == "kdv":
        return (ex.stepper.KortewegDeVries(1, 20.0, N, 1e-3, order=order),
                ost.KortewegDeVries(1, 20.0, N, 1e-3, order=order, dtype=np.float32), 20.0)
    return (ex.stepper.Advection(1, 2.0, N, 0.01, velocity=0.8),
            ost.Advection(1, 2.0, N, 0.01, velocity=0.8, dtype=np.float32), 2.0)


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("name,order", [("burgers", 2), ("burgers", 1), ("ks_cons", 2), ("ks", 2), ("ks", 4),
                                        ("ks_cons", 3), ("kdv", 2), ("advection", 0)])
def test_mixed_radix_parity(device, name, order, N):
    import torch

    print(f"[case] {name} order {order} N={N}", flush=True)
    gpu, cpu, L = _pair(name, N, order)
    u0, _ = ics(6, 1, L, N, 5, seed=48 + N)
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


@pytest.mark.parametrize("N", SIZES)
def test_mixed_radix_burgers_ensemble_and_resident(device, N):
    """Per-trajectory viscosity and the Fourier-resident repeat at N = 48/50."""
    import torch

    print(f"[case] burgers ensemble N={N}", flush=True)
    nus = np.linspace(0.01, 0.05, 8)
    gpu = ex.ensemble(lambda nu: ex.stepper.Burgers(1, 1.0, N, 1e-3, diffusivity=float(nu)), nus)
    cpu = ost.Ensemble([ost.Burgers(1, 1.0, N, 1e-3, diffusivity=float(nu), dtype=np.float32) for nu in nus],
                       dtype=np.float32)
    u0, _ = ics(8, 1, 1.0, N, 5, seed=7 + N)
    x = torch.tensor(u0, device=device)
    uc = u0
    for _ in range(10):
        uc = cpu.step_batch(uc)
    got = gpu.rollout_batch(x, 10).cpu().numpy()[:, 9]
    rep = ex.RepeatedStepper(ex.stepper.Burgers(1, 1.0, N, 1e-3, diffusivity=float(nus[0])), 10)
    got_rep = rep.step(x[:1]).cpu().numpy()
    e, e_rep = max_rel_l2(got, uc), max_rel_l2(got_rep, uc[:1])
    print(f"  10 steps: ensemble {e:.2e}, resident repeat {e_rep:.2e}", flush=True)
    assert e <= RTOL_STEP and e_rep <= RTOL_STEP


@pytest.mark.parametrize("N", SIZES)
def test_mixed_radix_transforms(device, N):
    import torch

    print(f"[case] fft N={N}", flush=True)
    rng = np.random.default_rng(N)
    u = rng.standard_normal((3, 2, N)).astype(np.float32)
    uh = ex.spectral.fft(torch.tensor(u, device=device), num_spatial_dims=1)
    want = np.fft.rfft(u.astype(np.float64), axis=-1)
    got = uh.cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    back = ex.spectral.ifft(uh, num_spatial_dims=1, num_points=N).cpu().numpy()
    print(f"  rfft err {err:.2e}, roundtrip {np.abs(back - u).max():.2e}", flush=True)
    assert err <= 1e-6
    np.testing.assert_allclose(back, u, atol=1e-5)


def test_difficulty_steppers_default_resolution(device):
    import torch

    g = ex.stepper.generic
    for st in (g.DifficultyConvectionStepper(), g.DifficultyGradientNormStepper(), g.DifficultyPolynomialStepper(),
               g.DifficultyNonlinearStepper(), g.DifficultyLinearStepper(), g.DifficultyLinearStepperSimple()):
        assert st.num_points == 48
        u0, _ = ics(2, 1, 1.0, 48, 3, seed=99)
        out = st.rollout_batch(torch.tensor(u0, device=device), 5).cpu().numpy()
        print(f"[case] {type(st).__name__} default N=48: finite {np.isfinite(out).all()}", flush=True)
        assert np.isfinite(out).all(), type(st).__name__
