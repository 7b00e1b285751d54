"""Two-channel convection and the general nonlinearity in 2-D on the GPU.

Burgers in 2-D (velocity (u, v), both forms of nonlin_fun/_convection.py:
104-171; stepper/_burgers.py), Kuramoto-Sivashinsky in conservative form in
2-D (two channels), single-channel conservative Burgers and
GeneralNonlinearStepper in 2-D (nonlin_fun/_general_nonlinear.py:89-119)
against the FP32 oracle at steps 1 and 10, through single steps and the
rollout, for ETDRK orders 1-4 and per-trajectory coefficients.
"""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("N", [64, 128])
@pytest.mark.parametrize("conservative,order", [(False, 2), (True, 2), (False, 1), (True, 3), (False, 4)])
def test_burgers_2d_parity(device, conservative, order, N):
    print(f"[case] 2-D Burgers conservative={conservative} order {order} N={N}", flush=True)
    kw = dict(diffusivity=0.01, conservative=conservative, order=order)
    gpu = ex.stepper.Burgers(2, 1.0, N, 1e-3, **kw)
    cpu = ost.Burgers(2, 1.0, N, 1e-3, dtype=np.float32, **kw)
    u0, _ = ics(3, 2, 1.0, N, 4, seed=71, channels=2)
    _check(gpu, cpu, u0, device, "burgers2d")


def test_burgers_2d_ensemble(device):
    """Per-trajectory viscosity with two channels per trajectory."""
    print("[case] 2-D Burgers ensemble", flush=True)
    nus = [0.005, 0.01, 0.02, 0.04]
    gpu = ex.ensemble(lambda nu: ex.stepper.Burgers(2, 1.0, 64, 1e-3, diffusivity=nu), nus)
    cpu = ost.Ensemble([ost.Burgers(2, 1.0, 64, 1e-3, diffusivity=nu, dtype=np.float32) for nu in nus],
                       dtype=np.float32)
    u0, _ = ics(4, 2, 1.0, 64, 4, seed=72, channels=2)
    _check(gpu, cpu, u0, device, "ensemble")


def test_ks_conservative_2d_parity(device):
    print("[case] 2-D KS conservative", flush=True)
    gpu = ex.stepper.KuramotoSivashinskyConservative(2, 30.0, 64, 0.05)
    cpu = ost.KuramotoSivashinskyConservative(2, 30.0, 64, 0.05, dtype=np.float32)
    u0, _ = ics(2, 2, 30.0, 64, 3, seed=73, channels=2)
    _check(gpu, cpu, 0.3 * u0, device, "ks_cons2d")


@pytest.mark.parametrize("order", [2, 4])
def test_single_channel_burgers_and_general_2d(device, order):
    print(f"[case] 2-D single-channel Burgers / general order {order}", flush=True)
    kw = dict(diffusivity=0.01, single_channel=True, conservative=True, order=order)
    gpu = ex.stepper.Burgers(2, 1.0, 64, 1e-3, **kw)
    cpu = ost.Burgers(2, 1.0, 64, 1e-3, dtype=np.float32, **kw)
    u0, _ = ics(3, 2, 1.0, 64, 4, seed=74)
    _check(gpu, cpu, u0, device, "burgers_single")
    lin, nonlin = (0.0, 0.0, 0.01), (0.3, -1.0, -0.5)
    gpu = ex.stepper.generic.GeneralNonlinearStepper(2, 2.0, 64, 1e-3, linear_coefficients=lin,
                                                     nonlinear_coefficients=nonlin, order=order)
    cpu = ost.GeneralStepper(2, 2.0, 64, 1e-3, linear_coefficients=lin, nonlinear_scales=nonlin, order=order,
                             dtype=np.float32)
    u0, _ = ics(3, 2, 2.0, 64, 4, seed=75)
    _check(gpu, cpu, u0, device, "general")


def _broadband_2d(B, C, N, seed):
    """White-noise fields (every mode populated, Nyquist included)."""
    rng = np.random.default_rng(seed)
    return (0.1 * rng.standard_normal((B, C, N, N))).astype(np.float32)


@pytest.mark.parametrize("name,order", [("burgers", 2), ("burgers", 4), ("kdv", 2), ("kdv", 3), ("advection", 0),
                                        ("ks", 1), ("kolmogorov", 2)])
def test_step_fourier_2d(device, name, order):
    """2-D step_fourier (_base_stepper.py:222-239, etdrk/_base_etdrk.py:64-80)
    runs on the spectra themselves: against the oracle's step_fourier on the
    spectra of broadband fields, and on spectra whose DC and Nyquist bins
    carry imaginary parts (which a physical round trip would drop), for real
    and complex L, ETDRK orders 0-4; in place as well."""
    import torch

    from oracle import spectral as osp

    print(f"[case] 2-D step_fourier {name} order {order}", flush=True)
    N = 64
    if name == "burgers":
        kw = dict(diffusivity=0.01, order=order)
        gpu, cpu, C = ex.stepper.Burgers(2, 1.0, N, 1e-3, **kw), ost.Burgers(2, 1.0, N, 1e-3, dtype=np.float32, **kw), 2
    elif name == "kdv":
        gpu = ex.stepper.KortewegDeVries(2, 20.0, N, 1e-3, order=order)
        cpu, C = ost.KortewegDeVries(2, 20.0, N, 1e-3, order=order, dtype=np.float32), 2
    elif name == "advection":
        gpu = ex.stepper.Advection(2, 1.0, N, 1e-3, velocity=(0.7, -0.4))
        cpu, C = ost.Advection(2, 1.0, N, 1e-3, velocity=(0.7, -0.4), dtype=np.float32), 1
    elif name == "ks":
        gpu = ex.stepper.KuramotoSivashinsky(2, 30.0, N, 0.05, order=order)
        cpu, C = ost.KuramotoSivashinsky(2, 30.0, N, 0.05, order=order, dtype=np.float32), 1
    else:
        gpu = ex.stepper.KolmogorovFlowVorticity(2, 2 * np.pi, N, 0.01, order=order)
        cpu, C = ost.KolmogorovFlowVorticity(2, 2 * np.pi, N, 0.01, order=order, dtype=np.float32), 1
    u0 = _broadband_2d(3, C, N, seed=76)
    uh = osp.fft(u0, num_spatial_dims=2).astype(np.complex64)
    rng = np.random.default_rng(77)
    uh2 = uh.copy()
    for ky in (0, N // 2):  # imaginary parts on the DC and Nyquist columns
        uh2[..., ky] += 1j * (0.5 * np.abs(uh).mean() * rng.standard_normal(uh2[..., ky].shape)).astype(np.float32)
    for spec in (uh, uh2):
        want = cpu.step_fourier(spec)
        x = torch.tensor(spec, device=device)
        got = gpu.step_fourier(x).cpu().numpy()
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"  max rel {err:.2e}", flush=True)
        assert err <= 1e-5
    # the imaginary DC/Nyquist parts matter: the physical step of irfftn(uh2) differs
    phys = osp.fft(cpu.step_batch(osp.ifft(uh2, num_spatial_dims=2, num_points=N).astype(np.float32)),
                   num_spatial_dims=2)
    assert np.abs(phys - cpu.step_fourier(uh2)).max() > 1e-3 * np.abs(want).max()
    # in place through the C ABI (uh_in == uh_out)
    from exponax_amd import _lib

    xin = torch.tensor(uh2, device=device)
    n = int(np.prod(uh2.shape[1:]))
    _lib.check(_lib.lib().exmi_step_fourier(gpu.plan(3).handle, xin.data_ptr(), n, xin.data_ptr(), n,
                                            _lib.stream_ptr()), "exmi_step_fourier")
    np.testing.assert_array_equal(xin.cpu().numpy(), got)


@pytest.mark.parametrize("name", ["ks", "kdv", "advection"])
def test_repeated_stepper_2d_fourier_substeps(device, name):
    """2-D RepeatedStepper sub-steps in Fourier space like the reference
    (_repeated_stepper.py:56-99): against the oracle's fft -> k x
    step_fourier -> ifft on broadband fields, for a real L (KS) and complex
    ones (KdV, advection), where physical sub-steps would drop the imaginary
    Nyquist parts."""
    import torch

    from oracle import spectral as osp

    print(f"[case] 2-D RepeatedStepper {name}", flush=True)
    N, k = 64, 5
    if name == "ks":
        gpu, cpu, C = (ex.stepper.KuramotoSivashinsky(2, 30.0, N, 0.05),
                       ost.KuramotoSivashinsky(2, 30.0, N, 0.05, dtype=np.float32), 1)
    elif name == "kdv":
        gpu, cpu, C = (ex.stepper.KortewegDeVries(2, 20.0, N, 1e-3),
                       ost.KortewegDeVries(2, 20.0, N, 1e-3, dtype=np.float32), 2)
    else:
        gpu = ex.stepper.Advection(2, 1.0, N, 0.01, velocity=(0.7, -0.4))
        cpu, C = ost.Advection(2, 1.0, N, 0.01, velocity=(0.7, -0.4), dtype=np.float32), 1
    u0 = _broadband_2d(2, C, N, seed=78)
    got = ex.RepeatedStepper(gpu, k).step(torch.tensor(u0, device=device)).cpu().numpy()
    uh = osp.fft(u0, num_spatial_dims=2).astype(np.complex64)
    for _ in range(k):
        uh = cpu.step_fourier(uh)
    want = osp.ifft(uh, num_spatial_dims=2, num_points=N).astype(np.float32)
    err = max_rel_l2(got, want)
    print(f"  rel-L2 vs oracle transform-once sub-steps {err:.2e}", flush=True)
    assert err <= RTOL_STEP


def test_nonlinear_fun_2d(device):
    """The nonlinear functions on 2-D spectra (nonlin_fun/*.py __call__)
    against the oracle: two-channel convection, gradient norm, vorticity
    convection with the Kolmogorov forcing, polynomial, and the two-channel
    convection of a complex-L plan (KdV)."""
    import torch

    from oracle import spectral as osp

    print("[case] 2-D nonlinear functions", flush=True)
    cases = [
        (ex.stepper.Burgers(2, 1.0, 64, 1e-3), ost.Burgers(2, 1.0, 64, 1e-3, dtype=np.float32), 1.0, 2),
        (ex.stepper.KuramotoSivashinsky(2, 30.0, 64, 0.05),
         ost.KuramotoSivashinsky(2, 30.0, 64, 0.05, dtype=np.float32), 30.0, 1),
        (ex.stepper.KolmogorovFlowVorticity(2, 2 * np.pi, 64, 0.01),
         ost.KolmogorovFlowVorticity(2, 2 * np.pi, 64, 0.01, dtype=np.float32), 2 * np.pi, 1),
        (ex.stepper.reaction.AllenCahn(2, 1.0, 64, 1e-3), ost.AllenCahn(2, 1.0, 64, 1e-3, dtype=np.float32), 1.0, 1),
        (ex.stepper.KortewegDeVries(2, 20.0, 64, 1e-3),
         ost.KortewegDeVries(2, 20.0, 64, 1e-3, dtype=np.float32), 20.0, 2),
    ]
    for gpu, cpu, L, C in cases:
        u0, _ = ics(3, 2, L, 64, 4, seed=77, channels=C)
        uh = osp.fft(u0, num_spatial_dims=2).astype(np.complex64)
        got = gpu.nonlinear_fun(torch.tensor(uh, device=device)).cpu().numpy()
        want = cpu.nonlinear_fun(uh)
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"  {type(gpu).__name__}: max rel {err:.2e}", flush=True)
        assert err <= 1e-5


@pytest.mark.parametrize("order", [1, 2, 4])
@pytest.mark.parametrize("N", [64, 128])
def test_single_channel_nonconservative_burgers_2d(device, order, N):
    """Single-channel non-conservative convection -b F[u (d_x u + d_y u)]
    (nonlin_fun/_convection.py:204-231, Burgers(single_channel=True,
    conservative=False)): steps and rollout against the FP32 oracle, and
    the nonlinear function itself."""
    import torch

    from oracle import spectral as osp

    print(f"[case] 2-D single-channel non-conservative Burgers order {order} N={N}", flush=True)
    kw = dict(diffusivity=0.01, single_channel=True, conservative=False, order=order)
    gpu = ex.stepper.Burgers(2, 1.0, N, 1e-3, **kw)
    cpu = ost.Burgers(2, 1.0, N, 1e-3, dtype=np.float32, **kw)
    u0, _ = ics(3, 2, 1.0, N, 4, seed=75)
    _check(gpu, cpu, u0, device, "burgers2d_single_nc")
    uh = osp.fft(u0, num_spatial_dims=2).astype(np.complex64)
    got = gpu.nonlinear_fun(torch.tensor(uh, device=device)).cpu().numpy()
    want = cpu.nonlinear_fun(uh)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  N(u_hat) max rel {err:.2e}", flush=True)
    assert err <= 1e-5
