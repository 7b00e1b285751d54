"""GPU path against the committed golden fixtures (tests/golden/oracle_*.npz,
written by tests/golden/make_golden.py): seeded inputs of the benchmark
workloads at reduced batch with the oracle's FP32 and FP64 states after 1
and 10 steps. The gate is rel-L2 <= 1e-5 per trajectory against FP32 (the
stated parity bar); the FP64 distance is held to the same bound."""
import os

import numpy as np
import pytest

import exponax_amd as ex
from tests.helpers import max_rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GATE = 1e-5


def _load(name):
    with np.load(os.path.join(GOLDEN, f"oracle_{name}.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _stepper(name, fx):
    if name == "ks_cons_1d":
        return ex.stepper.KuramotoSivashinskyConservative(1, 100.0, 256, 0.1)
    if name == "ks_1d":
        return ex.stepper.KuramotoSivashinsky(1, 60.0, 128, 0.05)
    return ex.ensemble(lambda nu: ex.stepper.Burgers(1, 1.0, 1024, 1e-3, diffusivity=float(nu)), fx["nu"])


@pytest.mark.parametrize("name", ["ks_cons_1d", "ks_1d", "burgers_1d"])
@pytest.mark.parametrize("path", ["rollout", "steps"])
def test_golden_fixtures(device, name, path):
    import torch

    print(f"[case] golden {name} via {path}", flush=True)
    fx = _load(name)
    st = _stepper(name, fx)
    x = torch.tensor(fx["u0"].astype(np.float32), device=device)
    if path == "rollout":  # the multi-step kernels (autotuned strategy)
        traj = st.rollout_batch(x, 10).cpu().numpy()
        got = {1: traj[:, 0], 10: traj[:, 9]}
    else:  # one launch per step
        got, u = {}, x
        for s in range(1, 11):
            u = st.step(u)
            if s in (1, 10):
                got[s] = u.cpu().numpy()
    for s in (1, 10):
        e32 = max_rel_l2(got[s], fx[f"f32_step{s}"])
        e64 = max_rel_l2(got[s], fx[f"f64_step{s}"])
        print(f"  step {s}: rel-L2 vs FP32 {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
        assert e32 <= GATE and e64 <= GATE, (s, e32, e64)
