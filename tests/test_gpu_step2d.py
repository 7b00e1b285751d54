"""GPU parity of the 2-D ETDRK2 path (row/column kernels) against the FP32
oracle: Kuramoto-Sivashinsky 2-D (gradient norm,
stepper/_kuramoto_sivashinsky.py:8-168) and Kolmogorov / Navier-Stokes
vorticity (stepper/_navier_stokes.py:13-327) with per-instance Reynolds
numbers, one and ten steps, rel-L2 <= 1e-5."""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


def _run(gpu_step, cpu_step, u0, steps=(1, 10)):
    ug, uc = u0, u0
    for s in range(1, max(steps) + 1):
        ug = gpu_step(ug)
        uc = cpu_step(uc)
        if s in steps:
            assert np.isfinite(ug).all()
            err = max_rel_l2(ug, uc)
            print(f"  step {s}: max rel-L2 {err:.2e}", flush=True)
            assert err <= RTOL_STEP


@pytest.mark.parametrize("N", [32, 64, 128])
def test_ks2d_parity(device, N):
    print(f"[case] ks2d N={N}", flush=True)
    L, dt = 60.0 * N / 128, 0.05
    gpu = ex.stepper.KuramotoSivashinsky(2, L, N, dt)
    cpu = ost.KuramotoSivashinsky(2, L, N, dt, dtype=np.float32)
    u0, _ = ics(3, 2, L, N, 5, seed=4)
    _run(ex.vmap(gpu), cpu.step_batch, u0)


@pytest.mark.parametrize("N", [32, 64, 128, 256])
def test_kolmogorov_ensemble_parity(device, N):
    print(f"[case] kolmogorov N={N}", flush=True)
    rng = np.random.default_rng(3)
    B = 3
    res = rng.uniform(100, 1000, B)
    u0, _ = ics(B, 2, 2 * np.pi, N, 6, seed=3)
    gpu = ex.ensemble(lambda re: ex.stepper.KolmogorovFlowVorticity(2, 2 * np.pi, N, 0.01, diffusivity=1.0 / re),
                      res)
    cpu = ost.Ensemble([ost.KolmogorovFlowVorticity(2, 2 * np.pi, N, 0.01, diffusivity=1.0 / re, dtype=np.float32)
                        for re in res], dtype=np.float32)
    _run(gpu, cpu.step_batch, u0)


def test_navier_stokes_parity(device):
    N = 64
    gpu = ex.stepper.NavierStokesVorticity(2, 1.0, N, 0.001, diffusivity=1e-3)
    cpu = ost.NavierStokesVorticity(2, 1.0, N, 0.001, diffusivity=1e-3, dtype=np.float32)
    u0, _ = ics(2, 2, 1.0, N, 4, seed=9)
    _run(ex.vmap(gpu), cpu.step_batch, u0)


def test_rollout2d_equals_steps(device):
    import torch

    N = 64
    gpu = ex.stepper.KuramotoSivashinsky(2, 30.0, N, 0.05)
    u0, _ = ics(2, 2, 30.0, N, 5, seed=7)
    x = torch.tensor(u0, device=device)
    tr = gpu.rollout_batch(x, 5, include_init=True).cpu().numpy()
    u = x
    for t in range(1, 6):
        u = gpu.step(u)
        assert max_rel_l2(tr[:, t], u.cpu().numpy()) <= 1e-6
    last = gpu.repeat_batch(x, 5).cpu().numpy()
    assert max_rel_l2(last, tr[:, -1]) <= 1e-6


@pytest.mark.parametrize("order", [0, 1, 3, 4])
@pytest.mark.parametrize("kind", ["ks", "kolmogorov"])
@pytest.mark.parametrize("N", [32, 128, 256])
def test_2d_etdrk_orders_parity(device, kind, order, N):
    """2-D ETDRK0 (E u, etdrk/_etdrk_0.py:30-34), ETDRK1 (E u + c1 N(u),
    etdrk/_etdrk_1.py:77-82), ETDRK3 (etdrk/_etdrk_3.py:191-212) and ETDRK4
    (etdrk/_etdrk_4.py:198-224) against the FP32 oracle, 1 and 10 steps."""
    print(f"[case] 2-D {kind} order={order} N={N}", flush=True)
    if kind == "ks":
        L = 30.0
        gpu = ex.stepper.KuramotoSivashinsky(2, L, N, 0.02, order=order)
        cpu = ost.KuramotoSivashinsky(2, L, N, 0.02, order=order, dtype=np.float32)
    else:
        L = 2 * np.pi
        gpu = ex.stepper.KolmogorovFlowVorticity(2, L, N, 0.002, diffusivity=0.01, order=order)
        cpu = ost.KolmogorovFlowVorticity(2, L, N, 0.002, diffusivity=0.01, order=order, dtype=np.float32)
    u0, _ = ics(2, 2, L, N, 5, seed=17)
    _run(ex.vmap(gpu), cpu.step_batch, u0)


def test_2d_etdrk0_rollout_exact(device):
    """ETDRK0 on a linear problem: the 2-D rollout equals exp(L t) applied
    to the IC in closed form (float32 rounding)."""
    import torch

    N, L, dt, n = 64, 20.0, 0.05, 20
    gpu = ex.stepper.KuramotoSivashinsky(2, L, N, dt, order=0)
    u0, _ = ics(2, 2, L, N, 4, seed=18)
    tr = gpu.rollout_batch(torch.tensor(u0, device=device), n).cpu().numpy()
    from oracle import spectral as osp

    lin = gpu.linear_operator.astype(np.complex128)
    uh = osp.fft(u0.astype(np.float64), num_spatial_dims=2)
    ref = osp.ifft(np.exp(lin * dt * n) * uh, num_spatial_dims=2, num_points=N)
    assert max_rel_l2(tr[:, -1], ref) <= 1e-5


@pytest.mark.parametrize("order", [3, 4])
def test_2d_high_order_rollout_equals_steps(device, order):
    import torch

    N = 64
    gpu = ex.stepper.KolmogorovFlowVorticity(2, 2 * np.pi, N, 0.002, diffusivity=0.01, order=order)
    u0, _ = ics(2, 2, 2 * np.pi, N, 5, seed=19)
    x = torch.tensor(u0, device=device)
    tr = gpu.rollout_batch(x, 4).cpu().numpy()
    u = x
    for t in range(4):
        u = gpu.step(u)
        assert max_rel_l2(tr[:, t], u.cpu().numpy()) <= 1e-6


def test_2d_rollout_resume_equals_continuous(device):
    import torch

    gpu = ex.stepper.KuramotoSivashinsky(2, 30.0, 64, 0.05)
    u0, _ = ics(2, 2, 30.0, 64, 5, seed=20)
    x = torch.tensor(u0, device=device)
    full = gpu.rollout_batch(x, 40).cpu().numpy()
    a = gpu.rollout_batch(x, 20)
    b = gpu.rollout_batch(a[:, -1].contiguous(), 20)
    got = torch.cat([a, b], dim=1).cpu().numpy()
    assert max_rel_l2(got[:, -1], full[:, -1]) <= 1e-6
