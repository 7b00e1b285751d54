"""GPU parity of the corrector network kernels (exmi_cnn_forward /
exmi_cnn_backward) against the oracle's bfloat16 model (oracle/cnn.py).

That model is pinned to torch's float64 conv1d/autograd in
tests/test_oracle_cnn.py. Tolerances are relative L2:
- forward correction f(x): 1e-3;
- parameter and input gradients: 5e-3.
Both sides round the same quantities to bfloat16. What remains is the float32
summation order, and the rare bfloat16 rounding it flips.

Sample counts are odd on purpose, so the per-wavefront sample loops and the
weight-gradient partial sums get ragged ends.
"""
import numpy as np
import pytest

from oracle import cnn as ocnn

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _params(seed):
    rng = np.random.default_rng(seed)
    p = ocnn.init_params(seed, scale_last=1.0)
    return (p + 0.02 * rng.standard_normal(p.shape)).astype(np.float32)


@pytest.mark.parametrize("N", [32, 64, 128])
@pytest.mark.parametrize("count", [1, 37])
def test_forward_parity(device, N, count):
    import torch

    from exponax_amd import corrector as cr

    print(f"[case] cnn forward N={N} count={count}", flush=True)
    p = _params(N + count)
    x = np.random.default_rng(1).standard_normal((count, 1, N)).astype(np.float32)
    pt = torch.tensor(p, device="cuda")
    y = cr.cnn_forward(pt, torch.tensor(x, device="cuda")).cpu().numpy()
    ref, _ = ocnn.forward(p, x, bf16_model=True)
    ref64, _ = ocnn.forward(p, x, bf16_model=False)
    err = _rel(y - x, ref - x)
    err64 = _rel(y - x, ref64 - x)
    print(f"  f(x) rel-L2 vs bf16 oracle {err:.2e}, vs float64 network {err64:.2e}", flush=True)
    assert np.isfinite(y).all()
    assert err <= 1e-3
    # with a workspace the same outputs come back
    ws = cr.workspace_for(N, count, "cuda")
    y2 = cr.cnn_forward(pt, torch.tensor(x, device="cuda"), workspace=ws).cpu().numpy()
    np.testing.assert_array_equal(y, y2)


@pytest.mark.parametrize("N", [32, 64, 128])
@pytest.mark.parametrize("count", [3, 301])
def test_backward_parity(device, N, count):
    import torch

    from exponax_amd import corrector as cr

    print(f"[case] cnn backward N={N} count={count}", flush=True)
    p = _params(7 * N + count)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((count, 1, N)).astype(np.float32)
    dy = rng.standard_normal((count, 1, N)).astype(np.float32)
    pt = torch.tensor(p, device="cuda", requires_grad=True)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    y = cr._CnnFn.apply(xt, pt)
    (y * torch.tensor(dy, device="cuda")).sum().backward()
    gp, gx = pt.grad.cpu().numpy(), xt.grad.cpu().numpy()
    rp, rx = ocnn.backward(p, x, dy, bf16_model=True)
    ep, ex_ = _rel(gp, rp), _rel(gx - dy, rx - dy)
    print(f"  d params rel-L2 {ep:.2e}, d x (network part) rel-L2 {ex_:.2e}", flush=True)
    assert ep <= 5e-3 and ex_ <= 5e-3
    # per-layer check (a layout slip in one layer cannot hide in the total)
    o = 0
    for ws_, bs_ in ocnn.SHAPES:
        nw, nb = int(np.prod(ws_)), int(np.prod(bs_))
        assert _rel(gp[o:o + nw], rp[o:o + nw]) <= 1e-2
        assert _rel(gp[o + nw:o + nw + nb], rp[o + nw:o + nw + nb]) <= 1e-2
        o += nw + nb


def test_backward_deterministic(device):
    import torch

    from exponax_amd import corrector as cr

    p = torch.tensor(_params(5), device="cuda", requires_grad=True)
    x = torch.randn(513, 1, 64, device="cuda")
    grads = []
    for _ in range(2):
        p.grad = None
        (cr._CnnFn.apply(x, p) ** 2).sum().backward()
        grads.append(p.grad.clone())
    assert torch.equal(grads[0], grads[1])


def _ks_pair():
    import exponax_amd as ex

    fine = ex.stepper.KuramotoSivashinskyConservative(1, 100.0, 256, 0.1)
    coarse = ex.stepper.KuramotoSivashinskyConservative(1, 100.0, 64, 0.5)
    return fine, coarse


def test_training_beats_uncorrected_coarse(device):
    """Ablation: after a short GPU training run on GPU-generated fine data,
    the corrected coarse stepper (N = 64, dt = 0.5) is more accurate than
    the same coarse stepper without the network, 10 coarse steps (5 time
    units) ahead on held-out trajectories."""
    import torch

    import exponax_amd as ex
    from exponax_amd import corrector as cr

    print("[case] corrector ablation", flush=True)
    fine, coarse = _ks_pair()
    u0 = torch.tensor(ex.ic.truncated_fourier_series(256, 1, 100.0, 256, 5, 5), dtype=torch.float32,
                      device="cuda")
    data = cr.downsampled_trajectories(fine, u0, 60, 5, 64)
    cnn = cr.PeriodicCNN(seed=0)
    hist = cr.train(cnn, coarse, data, iters=800, batch=256, unroll=3, lr=2e-3, log_every=200)
    print("  loss history", hist, flush=True)
    assert hist[-1][1] < hist[0][1]
    ut = torch.tensor(ex.ic.truncated_fourier_series(64, 1, 100.0, 256, 5, 50), dtype=torch.float32,
                      device="cuda")
    ref = cr.downsampled_trajectories(fine, ut, 10, 5, 64)
    corr = cr.CorrectedStepper(coarse, cnn).rollout_batch(ref[:, 0].contiguous(), 10)
    plain = coarse.rollout_batch(ref[:, 0].contiguous(), 10)
    ec = ex.metrics.mean_nRMSE(corr[:, -1], ref[:, -1])
    ep = ex.metrics.mean_nRMSE(plain[:, -1], ref[:, -1])
    print(f"  5 time units ahead: corrected {ec:.3e}, uncorrected {ep:.3e}", flush=True)
    assert ec < 0.7 * ep


def test_fallback_recomputes_exactly_the_failed_trajectories(device):
    """Corrupted coarse initial states (NaN and 1e30, forced from the host)
    are flagged and recomputed by the fine stepper; the other trajectories
    are the plain corrected rollout, bit for bit."""
    import torch

    import exponax_amd as ex
    from exponax_amd import corrector as cr

    fine, coarse = _ks_pair()
    cnn = cr.PeriodicCNN(seed=1)
    cs = cr.CorrectedStepper(coarse, cnn)
    u0 = torch.tensor(ex.ic.truncated_fourier_series(12, 1, 100.0, 256, 5, 9), dtype=torch.float32, device="cuda")
    u0c = ex.map_between_resolutions(u0, 64)
    bad0 = u0c.clone()
    bad0[3, 0, 5] = float("nan")
    bad0[7, 0, 11] = 1e30
    tr, mask = cr.rollout_with_fallback(cs, fine, u0, 6, 5, u0_coarse=bad0)
    assert mask.cpu().tolist() == [i in (3, 7) for i in range(12)]
    assert torch.isfinite(tr).all()
    ref_fine = cr.downsampled_trajectories(fine, u0[[3, 7]].contiguous(), 6, 5, 64)[:, 1:]
    assert torch.equal(tr[[3, 7]], ref_fine)
    plain = cs.rollout_batch(u0c, 6)
    keep = [i for i in range(12) if i not in (3, 7)]
    assert torch.equal(tr[keep], plain[keep])


@pytest.mark.parametrize("N", [64, 128])
def test_corrected_rollout_matches_oracle(device, N):
    """exmi_corrected_rollout (u_{t+1} = cnn(step(u_t)), no reference
    counterpart; rollout semantics of _utils.py:92-186) against the oracle's
    float32 coarse KS-conservative step composed with the oracle's bfloat16
    network model, for 5 steps of 37 trajectories; also against the
    composition of the library's own step and network launches. The network
    part alone matches the oracle to 1e-3 relative (test_forward_parity);
    each GPU step, taken from the GPU's previous state, matches the oracle's
    step from that state to 5e-4 (the correction is a fraction of the
    state). Chained over the 5 steps the coarse dynamics amplify the rare
    bfloat16 rounding flips (1.1e-3 measured), so the chained distance is
    gated at 3e-3."""
    import torch

    import exponax_amd as ex
    from exponax_amd import corrector as cr
    from oracle import stepper as ost
    from tests.helpers import ics, max_rel_l2

    print(f"[case] corrected rollout N={N}", flush=True)
    L, DT, K, B = 100.0, 0.5, 5, 37
    p = ocnn.init_params(7, scale_last=0.3)
    coarse = ex.stepper.KuramotoSivashinskyConservative(1, L, N, DT)
    cnn = cr.PeriodicCNN(p)
    u0, _ = ics(B, 1, L, N, 5, seed=70)
    x = torch.tensor(u0, device=device)
    tr = cr.CorrectedStepper(coarse, cnn).rollout_batch(x, K, include_init=True).cpu().numpy()
    np.testing.assert_array_equal(tr[:, 0], u0)
    o32 = ost.KuramotoSivashinskyConservative(1, L, N, DT, dtype=np.float32)
    u = u0
    ug = x
    for s in range(1, K + 1):
        u = ocnn.forward(p, o32.step_batch(u), bf16_model=True)[0].astype(np.float32)
        one = ocnn.forward(p, o32.step_batch(tr[:, s - 1]), bf16_model=True)[0].astype(np.float32)
        ug = cnn(coarse.step(ug))
        e_one, e_or = max_rel_l2(tr[:, s], one), max_rel_l2(tr[:, s], u)
        e_lib = max_rel_l2(tr[:, s], ug.cpu().numpy())
        print(f"  step {s}: one step vs oracle {e_one:.2e}, chained vs oracle {e_or:.2e}, "
              f"vs step+network launches {e_lib:.2e}", flush=True)
        assert e_one <= 5e-4 and e_or <= 3e-3 and e_lib <= 1e-5


def test_fallback_with_an_ensemble_fine_stepper(device):
    """rollout_with_fallback with per-trajectory fine coefficients (a Burgers
    viscosity ensemble as the fine stepper): the failed trajectories are
    recomputed with their own members' operators."""
    import torch

    import exponax_amd as ex
    from exponax_amd import corrector as cr

    nus = [0.01, 0.02, 0.03, 0.04, 0.05, 0.06]
    fine = ex.ensemble(lambda nu: ex.stepper.Burgers(1, 1.0, 256, 1e-3, diffusivity=nu), nus)
    coarse = ex.stepper.Burgers(1, 1.0, 64, 5e-3, diffusivity=0.03)
    cs = cr.CorrectedStepper(coarse, cr.PeriodicCNN(seed=2))
    u0 = torch.tensor(ex.ic.truncated_fourier_series(6, 1, 1.0, 256, 4, 11), dtype=torch.float32, device="cuda")
    bad0 = ex.map_between_resolutions(u0, 64).clone()
    bad0[4, 0, 3] = float("nan")
    tr, mask = cr.rollout_with_fallback(cs, fine, u0, 4, 5, u0_coarse=bad0)
    assert mask.cpu().tolist() == [i == 4 for i in range(6)]
    member = fine.members[4]
    want = cr.downsampled_trajectories(member, u0[4:5].contiguous(), 4, 5, 64)[:, 1:]
    other = cr.downsampled_trajectories(fine.members[0], u0[4:5].contiguous(), 4, 5, 64)[:, 1:]
    err = float(((tr[4:5] - want).norm() / want.norm()).item())
    err_other = float(((other - want).norm() / want.norm()).item())
    print(f"  fallback vs its own member {err:.2e}, vs another member {err_other:.2e}", flush=True)
    assert err <= 1e-6 < err_other  # the rollout strategies of the two plans agree to rounding
    assert torch.isfinite(tr).all()
