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
        # the workgroup-resident rollout keeps the spectrum on chip: the
        # steps' float32 round trips differ from it by rounding
        assert max_rel_l2(tr[:, t], u.cpu().numpy()) <= 5e-6
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
    assert max_rel_l2(got[:, -1], full[:, -1]) <= 1e-5  # the resumed half restarts from the float32 state


def _pair2d(kind, N, B):
    if kind == "ks":
        L, dt = 60.0 * N / 128, 0.05
        return (ex.stepper.KuramotoSivashinsky(2, L, N, dt),
                ost.KuramotoSivashinsky(2, L, N, dt, dtype=np.float32), ics(B, 2, L, N, 5, seed=40 + N)[0])
    if kind == "kolmogorov":
        L, dt = 2 * np.pi, 0.01
        res = np.random.default_rng(41).uniform(100, 1000, B)
        gpu = ex.ensemble(lambda re: ex.stepper.KolmogorovFlowVorticity(2, L, N, dt, diffusivity=1.0 / re), res)
        cpu = ost.Ensemble([ost.KolmogorovFlowVorticity(2, L, N, dt, diffusivity=1.0 / re, dtype=np.float32)
                            for re in res], dtype=np.float32)
        return gpu, cpu, ics(B, 2, L, N, 6, seed=41 + N)[0]
    if kind == "burgers":
        kw = dict(diffusivity=0.01)
        return (ex.stepper.Burgers(2, 1.0, N, 1e-3, **kw), ost.Burgers(2, 1.0, N, 1e-3, dtype=np.float32, **kw),
                ics(B, 2, 1.0, N, 4, seed=42 + N, channels=2)[0])
    if kind == "general":
        lin, nonlin = (0.0, 0.0, 0.01), (0.3, -1.0, -0.5)
        return (ex.stepper.generic.GeneralNonlinearStepper(2, 2.0, N, 1e-3, linear_coefficients=lin,
                                                           nonlinear_coefficients=nonlin),
                ost.GeneralStepper(2, 2.0, N, 1e-3, linear_coefficients=lin, nonlinear_scales=nonlin,
                                   dtype=np.float32), ics(B, 2, 2.0, N, 4, seed=43 + N)[0])
    return (ex.stepper.reaction.AllenCahn(2, 1.0, N, 1e-3), ost.AllenCahn(2, 1.0, N, 1e-3, dtype=np.float32),
            ics(B, 2, 1.0, N, 4, seed=44 + N)[0])


@pytest.mark.parametrize("N", [32, 64, 128, 256])
@pytest.mark.parametrize("kind", ["ks", "kolmogorov", "burgers", "general", "allen_cahn"])
def test_resident_rollout_2d(device, kind, N):
    """The spectrum-resident 2-D rollouts, for plans whose per-step forward
    transform of the emitted state is an exact identity (real L, Nyquist
    row and column dealiased): the workgroup-resident kernel (strategy 1,
    N <= 128, gradient norm and vorticity) and the four-launch chain that
    carries S' in the column layout (strategy 2, every kind) and the same
    chain as one cooperative launch (strategy 3: N = 128 and 256, gradient
    norm and vorticity), against the per-step row/column launches (strategy -1: the float32 physical round
    trip every step) and the FP32 oracle at steps 1 and 10, with
    per-instance operator tables (Kolmogorov ensemble), two channels
    (Burgers) and shared tables; repeat equals the rollout's last state;
    include_init keeps u0."""
    import torch

    print(f"[case] spectrum-resident 2-D rollout {kind} N={N}", flush=True)
    B, n = 3, 12
    gpu, cpu, u0 = _pair2d(kind, N, B)
    C = u0.shape[1]
    x = torch.tensor(u0, device=device)
    plan = gpu.plan(B)
    outs = {}
    for strat in (-1, 1, 2, 3):
        try:
            plan.set_rollout_strategy(strat)
        except NotImplementedError:
            continue
        outs[strat] = gpu.rollout_batch(x, n, include_init=True).cpu().numpy()
        np.testing.assert_array_equal(outs[strat][:, 0], u0)
    assert 2 in outs and (1 in outs) == (N <= 128 and kind in ("ks", "kolmogorov"))
    assert (3 in outs) == (N in (128, 256) and kind in ("ks", "kolmogorov"))
    if 3 in outs:
        # the cooperative launch runs the chain's passes: the same arithmetic
        d = max_rel_l2(outs[3][:, 1:].reshape(B * n, C, N, N), outs[2][:, 1:].reshape(B * n, C, N, N))
        print(f"  strategy 3 vs 2: {d:.2e} (bit-identical: {np.array_equal(outs[3], outs[2])})", flush=True)
        assert d <= 1e-6
    ref = outs[-1]
    for strat, tr in outs.items():
        d = max_rel_l2(tr[:, 1:].reshape(B * n, C, N, N), ref[:, 1:].reshape(B * n, C, N, N))
        print(f"  strategy {strat} vs per-step launches: {d:.2e}", flush=True)
        assert d <= 5e-6
        u = u0
        for s in range(1, 11):
            u = cpu.step_batch(u)
            if s in (1, 10):
                e = max_rel_l2(tr[:, s], u)
                print(f"  strategy {strat} step {s}: rel-L2 vs oracle {e:.2e}", flush=True)
                assert e <= RTOL_STEP
        plan.set_rollout_strategy(strat)
        last = gpu.repeat_batch(x, n).cpu().numpy()
        np.testing.assert_array_equal(last, tr[:, -1])


def test_cooperative_rollout_capacity_and_group(device):
    """The cooperative 2-D launch needs every workgroup of its grid resident:
    a batch beyond what the device holds at once (9 workgroups per 128^2
    instance) does not offer it. A plan tuned to it runs the launch chain
    inside a group of several plans, with the same states."""
    import torch

    print("[case] cooperative rollout capacity and group", flush=True)
    st = ex.stepper.KuramotoSivashinsky(2, 60.0, 128, 0.05)
    small, big = st.plan(4), st.plan(512)
    small.set_rollout_strategy(3)
    with pytest.raises(NotImplementedError):
        big.set_rollout_strategy(3)
    u0 = torch.tensor(ics(4, 2, 60.0, 128, 5, seed=61)[0], device=device)
    alone = st.rollout_batch(u0, 12).cpu().numpy()
    other = ex.stepper.KuramotoSivashinsky(2, 60.0, 64, 0.05)
    v0 = torch.tensor(ics(2, 2, 60.0, 64, 5, seed=62)[0], device=device)
    for mode in (1, 2, 3):
        a, _ = ex.rollout_group([st, other], [u0, v0], 12, use_graph=mode)
        np.testing.assert_array_equal(a.cpu().numpy(), alone)
    assert small.rollout_strategy() == 3


def test_cooperative_barrier_give_up(device, monkeypatch):
    """A cooperative barrier that runs past its poll limit gives up: the
    launch ends normally and the states it has not emitted are NaN rather
    than stale data (forced from the host with a poll limit of 1, so the
    workgroups that wait on slower ones give up). The next launch starts
    from fresh counters: a plan with the default limit is unaffected."""
    import torch

    print("[case] cooperative barrier give-up", flush=True)
    u0 = torch.tensor(ics(4, 2, 60.0, 128, 5, seed=63)[0], device=device)
    monkeypatch.setenv("EXMI_COOP_SPIN_LIMIT", "1")
    bad = ex.stepper.KuramotoSivashinsky(2, 60.0, 128, 0.05)
    bad.plan(4).set_rollout_strategy(3)
    tr = bad.rollout_batch(u0, 12).cpu().numpy()
    assert not np.isfinite(tr).all()
    g = bad.rollout_vjp(u0, 12, g_final=u0, checkpoint_every=6).cpu().numpy()
    assert not np.isfinite(g).all()
    monkeypatch.delenv("EXMI_COOP_SPIN_LIMIT")
    good = ex.stepper.KuramotoSivashinsky(2, 60.0, 128, 0.05)
    good.plan(4).set_rollout_strategy(3)
    a = good.rollout_batch(u0, 12).cpu().numpy()
    good.plan(4).set_rollout_strategy(2)
    b = good.rollout_batch(u0, 12).cpu().numpy()
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(a, b)
