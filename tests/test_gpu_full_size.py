"""The benchmark workloads at their full sizes.

Oracle parity on a subset of each batch at steps 1 and 10 (the full batch
runs on the GPU, so batch-size dependent launch geometry is exercised), and
a size-independent property over the whole run length:
- Burgers 1-D, 4096 x N=1024, per-trajectory viscosity, 1000 steps: the
  spatial mean of every trajectory is invariant (the mode k = 0 has L = 0
  and a zero nonlinear term), so zero-mean initial states stay zero-mean on
  every trajectory that stays finite;
- Kolmogorov 2-D, 32 x 128^2 + 16 x 256^2 + 8 x 512^2: the mean vorticity
  obeys d<w>/dt = drag <w> (forcing and the divergence-free advection have
  no mean mode), so zero-mean initial fields stay zero-mean.
Trajectories that blow up do so in the reference algorithm too (the
lowest-viscosity Burgers members; Kolmogorov at dt = 0.01), and are left
out of the property checks.
"""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


def test_burgers_full_batch(device):
    import torch

    B, N, L, DT = 4096, 1024, 1.0, 1e-3
    print(f"[case] burgers full batch B={B} N={N}", flush=True)
    u0, rng = ics(B, 1, L, N, 8, seed=2)
    nus = 10.0 ** rng.uniform(-4.0, -2.0, B)
    gpu = ex.ensemble(lambda nu: ex.stepper.Burgers(1, L, N, DT, diffusivity=float(nu)), nus)
    x = torch.tensor(u0, device=device)
    traj = gpu.rollout_batch(x, 10).cpu().numpy()
    sub = np.arange(0, B, B // 32)
    cpu = ost.Ensemble([ost.Burgers(1, L, N, DT, diffusivity=float(nus[i]), dtype=np.float32) for i in sub],
                       dtype=np.float32)
    u = u0[sub]
    for s in range(1, 11):
        u = cpu.step_batch(u)
        if s in (1, 10):
            err = max_rel_l2(traj[sub, s - 1], u)
            print(f"  step {s}: max rel-L2 over {len(sub)} trajectories {err:.2e}", flush=True)
            assert err <= RTOL_STEP
    last = gpu.repeat_batch(x, 1000).cpu().numpy()
    finite = np.isfinite(last).all(axis=(1, 2))
    scale = np.abs(u0).max(axis=(1, 2))
    drift = np.abs(last.mean(axis=(1, 2)) - u0.mean(axis=(1, 2)))[finite] / scale[finite]
    print(f"  1000 steps: {finite.sum()} finite, max mean drift {drift.max():.2e}", flush=True)
    assert finite.sum() >= B * 0.9  # only the lowest-viscosity members blow up
    assert drift.max() <= 1e-5


def test_kolmogorov_full_mix(device):
    import torch

    sizes = ((128, 32), (256, 16), (512, 8))
    L, DT = 2 * np.pi, 0.01
    rng = np.random.default_rng(3)
    steppers, u0s, res_all = [], [], []
    for N, B in sizes:
        res = rng.uniform(100, 1000, B)
        res_all.append(res)
        u0, _ = ics(B, 2, L, N, 6, seed=3 + N)
        u0s.append(u0)
        steppers.append(ex.ensemble(lambda re, N=N: ex.stepper.KolmogorovFlowVorticity(
            2, L, N, DT, diffusivity=1.0 / re), res))
    print("[case] kolmogorov full mix", flush=True)
    xs = [torch.tensor(u, device=device) for u in u0s]
    trajs = ex.rollout_group(steppers, xs, 10, use_graph=2)
    for (N, B), res, u0, tr in zip(sizes, res_all, u0s, trajs):
        cpu = ost.KolmogorovFlowVorticity(2, L, N, DT, diffusivity=1.0 / res[0], dtype=np.float32)
        u = u0[:1]
        tr = tr.cpu().numpy()
        for s in range(1, 11):
            u = cpu.step_batch(u)
            if s in (1, 10):
                err = max_rel_l2(tr[:1, s - 1], u)
                print(f"  N={N} step {s}: rel-L2 {err:.2e}", flush=True)
                assert err <= RTOL_STEP
    finals = [st.repeat_batch(x, 100).cpu().numpy() for st, x in zip(steppers, xs)]
    for (N, B), u0, fin in zip(sizes, u0s, finals):
        finite = np.isfinite(fin).all(axis=(1, 2, 3))
        m = np.abs(fin.mean(axis=(1, 2, 3)))[finite] / np.abs(u0).max(axis=(1, 2, 3))[finite]
        print(f"  N={N}: {finite.sum()}/{B} finite after 100 steps, max |mean| {m.max() if m.size else 0:.2e}",
              flush=True)
        assert m.size == 0 or m.max() <= 1e-5
