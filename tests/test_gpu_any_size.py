"""Grid sizes outside the fused kernels' set (the reference accepts any N,
_spectral.py:614-721): odd sizes, primes, N = 100 (the learning notebook's
resolution), N = 513 (the 2-D Kolmogorov spectrum validation), in 1-D, 2-D
and 3-D. The runtime-size mixed-radix transforms against numpy (float64);
steppers against the FP32 oracle at steps 1 and 10 (rel-L2 <= 1e-5), for
ETDRK orders 0-4, real and complex linear operators, every nonlinear kind,
per-trajectory operator tables, step_fourier, the nonlinear function and
RepeatedStepper; the transform autotuner's choice (timed, persisted in
EXMI_TUNE_CACHE and reused) and parity under every candidate it times."""
import os

import numpy as np
import pytest

import exponax_amd as ex
from oracle import spectral as osp
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# prime sizes run one direct-DFT pass of their own length
FFT_SIZES_1D = [2, 3, 5, 7, 11, 12, 24, 97, 100, 210, 243, 343, 513, 1000, 1009, 2310, 4095, 17 * 19, 13 * 11 * 9]


@pytest.mark.parametrize("N", FFT_SIZES_1D)
def test_fft_any_size_1d_against_numpy(device, N):
    import torch

    print(f"[case] 1-D transforms N={N}", flush=True)
    rng = np.random.default_rng(N)
    x = rng.standard_normal((5, N)).astype(np.float32)
    xh = ex.spectral.fft(torch.tensor(x, device=device), num_spatial_dims=1).cpu().numpy()
    ref = np.fft.rfft(x.astype(np.float64), axis=-1)
    tol = 3e-6 * np.sqrt(max(osp_largest_factor(N), 8))
    assert xh.shape == ref.shape
    assert _rel(xh, ref) <= tol
    # inverse of an arbitrary half spectrum: Im of DC (and Nyquist) ignored as numpy does
    yh = (rng.standard_normal((5, N // 2 + 1)) + 1j * rng.standard_normal((5, N // 2 + 1))).astype(np.complex64)
    y = ex.spectral.ifft(torch.tensor(yh, device=device), num_spatial_dims=1, num_points=N).cpu().numpy()
    yref = np.fft.irfft(yh.astype(np.complex128), n=N, axis=-1)
    assert _rel(y, yref) <= tol


def osp_largest_factor(n):
    f, big = 2, 1
    while f * f <= n:
        while n % f == 0:
            big, n = max(big, f), n // f
        f += 1
    return max(big, n)


@pytest.mark.parametrize("D,N", [(2, 11), (2, 100), (2, 96), (2, 513), (2, 1024), (3, 12), (3, 21), (3, 50)])
def test_fft_any_size_nd_against_numpy(device, D, N):
    import torch

    print(f"[case] {D}-D transforms N={N}", flush=True)
    rng = np.random.default_rng(7 * N + D)
    x = rng.standard_normal((2,) + (N,) * D).astype(np.float32)
    xh = ex.spectral.fft(torch.tensor(x, device=device), num_spatial_dims=D).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64), axes=tuple(range(1, D + 1)))
    assert _rel(xh, ref) <= 2e-5
    back = ex.spectral.ifft(torch.tensor(xh, device=device), num_spatial_dims=D, num_points=N).cpu().numpy()
    assert _rel(back, x) <= 2e-5


def _pair1d(name, N, order):
    if name == "burgers":
        kw = dict(diffusivity=0.02, order=order)
        return ex.stepper.Burgers(1, 1.0, N, 1e-3, **kw), ost.Burgers(1, 1.0, N, 1e-3, dtype=np.float32, **kw), 1.0
    if name == "burgers_nc":
        kw = dict(diffusivity=0.02, order=order, conservative=False)
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
    if name == "allen_cahn":
        return (ex.stepper.reaction.AllenCahn(1, 1.0, N, 1e-3, order=order),
                ost.AllenCahn(1, 1.0, N, 1e-3, order=order, dtype=np.float32), 1.0)
    return (ex.stepper.Advection(1, 2.0, N, 0.01, velocity=0.8),
            ost.Advection(1, 2.0, N, 0.01, velocity=0.8, dtype=np.float32), 2.0)


def _check_steps(gpu, cpu, u0, device, label):
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


@pytest.mark.parametrize("N", [11, 100, 243])
@pytest.mark.parametrize("name,order", [("burgers", 2), ("burgers_nc", 1), ("ks_cons", 2), ("ks", 2), ("ks", 4),
                                        ("ks_cons", 3), ("kdv", 2), ("allen_cahn", 2), ("advection", 0)])
def test_any_size_1d_parity(device, name, order, N):
    print(f"[case] {name} order {order} N={N}", flush=True)
    gpu, cpu, L = _pair1d(name, N, order)
    u0, _ = ics(6, 1, L, N, min(5, (N - 1) // 2), seed=300 + N)
    _check_steps(gpu, cpu, u0, device, f"{name} N={N}")


def test_any_size_ensemble_and_coupled_1d(device):
    """Per-trajectory viscosity (lin_batch = B) and Gray-Scott (two coupled
    channels, per-channel L) at N = 100."""
    print("[case] burgers ensemble and gray-scott N=100", flush=True)
    N = 100
    nus = np.linspace(0.01, 0.05, 8)
    gpu = ex.ensemble(lambda nu: ex.stepper.Burgers(1, 1.0, N, 1e-3, diffusivity=float(nu)), nus)
    cpu = ost.Ensemble([ost.Burgers(1, 1.0, N, 1e-3, diffusivity=float(nu), dtype=np.float32) for nu in nus],
                       dtype=np.float32)
    u0, _ = ics(8, 1, 1.0, N, 5, seed=11)
    _check_steps(gpu, cpu, u0, device, "ensemble")
    gs = ex.stepper.reaction.GrayScott(1, 2.5, N, 1.0)
    gs_cpu = ost.GrayScott(1, 2.5, N, 1.0, dtype=np.float32)
    rng = np.random.default_rng(5)
    v0 = np.stack([1.0 - 0.3 * rng.random((3, N)), 0.3 * rng.random((3, N))], axis=1).astype(np.float32)
    _check_steps(gs, gs_cpu, v0, device, "gray-scott")


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
    if kind == "burgers_single_nc":
        kw = dict(diffusivity=0.01, single_channel=True, conservative=False)
        return (ex.stepper.Burgers(2, 1.0, N, 1e-3, **kw), ost.Burgers(2, 1.0, N, 1e-3, dtype=np.float32, **kw),
                ics(B, 2, 1.0, N, 4, seed=45 + N)[0])
    if kind == "general":
        lin, nonlin = (0.0, 0.0, 0.01), (0.3, -1.0, -0.5)
        return (ex.stepper.generic.GeneralNonlinearStepper(2, 2.0, N, 1e-3, linear_coefficients=lin,
                                                           nonlinear_coefficients=nonlin),
                ost.GeneralStepper(2, 2.0, N, 1e-3, linear_coefficients=lin, nonlinear_scales=nonlin,
                                   dtype=np.float32), ics(B, 2, 2.0, N, 4, seed=43 + N)[0])
    if kind == "dispersion":
        return (ex.stepper.Dispersion(2, 2.0, N, 1e-3, dispersivity=0.01),
                ost.Dispersion(2, 2.0, N, 1e-3, dispersivity=0.01, dtype=np.float32),
                ics(B, 2, 2.0, N, 4, seed=46 + N)[0])
    return (ex.stepper.reaction.AllenCahn(2, 1.0, N, 1e-3), ost.AllenCahn(2, 1.0, N, 1e-3, dtype=np.float32),
            ics(B, 2, 1.0, N, 4, seed=44 + N)[0])


@pytest.mark.parametrize("N", [33, 100])
@pytest.mark.parametrize("kind", ["ks", "kolmogorov", "burgers", "burgers_single_nc", "general", "dispersion",
                                  "allen_cahn"])
def test_any_size_2d_parity(device, kind, N):
    print(f"[case] 2-D {kind} N={N}", flush=True)
    gpu, cpu, u0 = _pair2d(kind, N, 3)
    _check_steps(gpu, cpu, u0, device, f"2-D {kind} N={N}")


def test_kolmogorov_513(device):
    """The 2-D Kolmogorov flow at the reference validation notebook's
    N = 513 (validation/validate_kolmogorov_2d_spectrum.ipynb): steps 1 and
    10 against the FP32 oracle, and its spectrum against the oracle's."""
    import torch

    print("[case] kolmogorov N=513", flush=True)
    gpu, cpu, u0 = _pair2d("kolmogorov", 513, 1)
    _check_steps(gpu, cpu, u0, device, "kolmogorov 513")
    x = torch.tensor(u0, device=device)
    spec = ex.get_spectrum(x[0]).cpu().numpy()
    ref = osp.get_spectrum(u0[0].astype(np.float64))
    assert spec.shape == ref.shape
    assert _rel(spec, ref) <= 1e-5


@pytest.mark.parametrize("N,name", [(12, "burgers"), (20, "ks"), (21, "ks")])
def test_any_size_3d_parity(device, N, name):
    print(f"[case] 3-D {name} N={N}", flush=True)
    if name == "burgers":
        gpu = ex.stepper.Burgers(3, 1.0, N, 1e-3, diffusivity=0.01)
        cpu = ost.Burgers(3, 1.0, N, 1e-3, diffusivity=0.01, dtype=np.float32)
        u0 = ics(2, 3, 1.0, N, 3, seed=50 + N, channels=3)[0]
    else:
        L = 20.0
        gpu = ex.stepper.KuramotoSivashinsky(3, L, N, 0.05)
        cpu = ost.KuramotoSivashinsky(3, L, N, 0.05, dtype=np.float32)
        u0 = ics(2, 3, L, N, 3, seed=51 + N)[0]
    _check_steps(gpu, cpu, u0, device, f"3-D {name} N={N}")


def test_any_size_spectral_entry_points(device):
    """step_fourier, the nonlinear function and RepeatedStepper at N = 100
    (1-D KdV, complex L) and N = 60 (2-D KS), against the oracle's
    fft -> k x step_fourier -> ifft path on a broadband state."""
    import torch

    print("[case] spectral entry points", flush=True)
    rng = np.random.default_rng(9)
    for D, N, gpu, cpu in [(1, 100, ex.stepper.KortewegDeVries(1, 20.0, 100, 1e-3),
                            ost.KortewegDeVries(1, 20.0, 100, 1e-3, dtype=np.float32)),
                           (2, 60, ex.stepper.KuramotoSivashinsky(2, 30.0, 60, 0.05),
                            ost.KuramotoSivashinsky(2, 30.0, 60, 0.05, dtype=np.float32))]:
        u = rng.standard_normal((2, 1) + (N,) * D).astype(np.float32) * 0.1
        uh = np.stack([osp.fft(v, num_spatial_dims=D) for v in u]).astype(np.complex64)
        got = gpu.step_fourier(torch.tensor(uh, device=device)).cpu().numpy()
        ref = np.stack([cpu.step_fourier(v) for v in uh])
        assert _rel(got, ref) <= 1e-5, (D, N, _rel(got, ref))
        nl = gpu.nonlinear_fun(torch.tensor(uh, device=device)).cpu().numpy()
        nl_ref = np.stack([cpu.nonlinear_fun(v) for v in uh])
        assert _rel(nl, nl_ref) <= 1e-5, (D, N, _rel(nl, nl_ref))
        rep = ex.RepeatedStepper(gpu, 5)
        got_r = rep.step(torch.tensor(u, device=device)).cpu().numpy()
        vh = uh.copy()
        for _ in range(5):
            vh = np.stack([cpu.step_fourier(v) for v in vh])
        ref_r = np.stack([osp.ifft(v, num_spatial_dims=D, num_points=N) for v in vh])
        assert max_rel_l2(got_r, ref_r) <= 1e-5


def test_transform_autotuner_choice_persisted_and_every_candidate(device, tmp_path):
    """The plan times the candidate radices / tiles when it is built, stores
    the choice in EXMI_TUNE_CACHE, and a second plan of the same shape reads
    it back; every candidate the tuner can pick gives the same parity."""
    import torch

    print("[case] transform autotuner", flush=True)
    cache = tmp_path / "tune.txt"
    old = os.environ.get("EXMI_TUNE_CACHE")
    os.environ["EXMI_TUNE_CACHE"] = str(cache)
    try:
        N = 100
        gpu, cpu, L = _pair1d("ks", N, 2)
        cfg = gpu.fft_config(batch=4)
        print(f"  chosen: {cfg}", flush=True)
        assert cfg is not None and cfg["source"] == "timed"
        assert int(np.prod(cfg["radices"])) == N
        lines = cache.read_text().split("\n")
        assert any("|gfft|D1|N100|I4 " in ln for ln in lines)
        again, _, _ = _pair1d("ks", N, 2)
        cfg2 = again.fft_config(batch=4)
        assert cfg2["source"] == "cache" and cfg2["code"] == cfg["code"]
        assert ex.stepper.Burgers(1, 1.0, 128, 1e-3).fft_config(batch=4) is None  # fused sizes have no knobs
        u0, _ = ics(4, 1, L, N, 5, seed=17)
        ref = u0
        for _ in range(10):
            ref = cpu.step_batch(ref)
        plan = gpu.plan(4)
        for code in (400000 + 1024, 800000 + 1024, 1600000 + 1024, 400000 + 2048, 800000 + 2048, 1600000 + 2048,
                     800000 + 512, 800000 + 4096, 200000 + 100, 10000000 + 100, 100000000 + 1600000 + 1024,
                     100000000 + 800000 + 2048):
            plan.set_fft_config(code)
            got = gpu.rollout_batch(torch.tensor(u0, device=device), 10).cpu().numpy()[:, -1]
            e = max_rel_l2(got, ref)
            print(f"  code {code}: radices {plan.fft_config()['radices']}, 10-step rel-L2 {e:.2e}", flush=True)
            assert e <= RTOL_STEP
    finally:
        if old is None:
            os.environ.pop("EXMI_TUNE_CACHE", None)
        else:
            os.environ["EXMI_TUNE_CACHE"] = old


def test_any_size_rollout_forms(device):
    """Rollouts of a multi-pass plan: include_init, the captured graph and
    direct launches agree, repeat equals the rollout's last state, a group
    mixes a runtime size with a fused one, and the failure-detecting rollout
    finds nothing on a stable run."""
    import torch

    print("[case] any-size rollout forms", flush=True)
    st = ex.stepper.KuramotoSivashinskyConservative(1, 30.0, 100, 0.05)
    u0, _ = ics(3, 1, 30.0, 100, 5, seed=81)
    x = torch.tensor(u0, device=device)
    a = st.rollout_batch(x, 12, include_init=True)
    b = st.rollout_batch(x, 12, use_graph=False)
    assert torch.equal(a[:, 0], x)
    assert max_rel_l2(a[:, 1:].cpu().numpy().reshape(3, -1), b.cpu().numpy().reshape(3, -1)) <= 1e-6
    r = st.repeat_batch(x, 12)
    assert max_rel_l2(r.cpu().numpy(), b[:, -1].cpu().numpy()) <= 1e-6
    traj, fb = st.rollout_checked(x, 12)
    assert (fb.cpu().numpy() == -1).all()
    fused = ex.stepper.KuramotoSivashinskyConservative(1, 30.0, 128, 0.05)
    u1, _ = ics(2, 1, 30.0, 128, 5, seed=82)
    y = torch.tensor(u1, device=device)
    for mode in (0, 1, 2, 3):
        ga, gb = ex.rollout_group([st, fused], [x, y], 12, use_graph=mode)
        assert max_rel_l2(ga.cpu().numpy().reshape(3, -1), b.cpu().numpy().reshape(3, -1)) <= 1e-6
        assert max_rel_l2(gb.cpu().numpy().reshape(2, -1),
                          fused.rollout_batch(y, 12).cpu().numpy().reshape(2, -1)) <= 1e-6


def test_any_size_noalloc_inverse_rejected(device):
    """The allocation-free inverse transform covers the fused sizes only: a
    runtime size is an error, not a silent allocation."""
    import ctypes

    import torch

    from exponax_amd import _lib

    print("[case] noalloc runtime size rejected", flush=True)
    L = _lib.lib()
    xh = torch.zeros(2, 51, dtype=torch.complex64, device=device)
    out = torch.zeros(2, 100, dtype=torch.float32, device=device)
    rc = L.exmi_irfft_noalloc(1, 100, 2, ctypes.c_void_p(xh.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                              ctypes.c_void_p(_lib.stream_ptr()))
    assert rc == _lib.EXMI_ERR_UNSUPPORTED


@pytest.mark.parametrize("case", ["burgers1d_101", "kolmogorov2d_100", "ks2d_96", "ns3d_32", "burgers3d_12"])
def test_multipass_spectrum_carrying_rollout(device, case):
    """Multi-pass plans with a real L (and, for an even N, the Nyquist modes
    dealiased): the rollout that keeps the spectrum from step to step and
    only inverse-transforms the states it emits (strategy 2) against the
    per-step round trip (-1) and the FP32 oracle at steps 1 and 10; repeat
    equals its last state and include_init keeps u0. A complex L offers only
    the per-step round trip."""
    import torch

    print(f"[case] multi-pass spectrum-carrying rollout {case}", flush=True)
    if case == "burgers1d_101":
        kw = dict(diffusivity=0.01)
        gpu, cpu = ex.stepper.Burgers(1, 1.0, 101, 1e-3, **kw), ost.Burgers(1, 1.0, 101, 1e-3, dtype=np.float32, **kw)
        u0 = ics(3, 1, 1.0, 101, 6, seed=91)[0]
    elif case == "kolmogorov2d_100":
        gpu, cpu, u0 = _pair2d("kolmogorov", 100, 2)
    elif case == "ks2d_96":
        gpu, cpu, u0 = _pair2d("ks", 96, 2)
    elif case == "ns3d_32":
        L, kw = 2 * np.pi, dict(diffusivity=0.02)
        gpu = ex.stepper.NavierStokesVelocity(3, L, 32, 0.01, **kw)
        cpu = ost.NavierStokesVelocity(3, L, 32, 0.01, dtype=np.float32, **kw)
        u0 = ics(2, 3, L, 32, 3, seed=121, channels=3)[0]
    else:
        kw = dict(diffusivity=0.01)
        gpu, cpu = ex.stepper.Burgers(3, 1.0, 12, 1e-3, **kw), ost.Burgers(3, 1.0, 12, 1e-3, dtype=np.float32, **kw)
        u0 = ics(2, 3, 1.0, 12, 3, seed=92, channels=3)[0]
    B, n = u0.shape[0], 12
    x = torch.tensor(u0, device=device)
    plan = gpu.plan(B)
    outs = {}
    for strat in (-1, 2):
        plan.set_rollout_strategy(strat)
        outs[strat] = gpu.rollout_batch(x, n, include_init=True).cpu().numpy()
        np.testing.assert_array_equal(outs[strat][:, 0], u0)
        last = gpu.repeat_batch(x, n).cpu().numpy()
        np.testing.assert_array_equal(last, outs[strat][:, -1])
    d = max(max_rel_l2(outs[2][:, s], outs[-1][:, s]) for s in range(1, n + 1))
    print(f"  carried vs per-step round trip, max over {n} steps: {d:.2e}", flush=True)
    u, errs = u0, []
    for s in range(1, 11):
        u = cpu.step_batch(u)
        if s in (1, 10):
            for strat, tr in outs.items():
                e = max_rel_l2(tr[:, s], u)
                print(f"  strategy {strat} step {s}: rel-L2 vs oracle {e:.2e}", flush=True)
                errs.append(e)
    # both are float32 computations of the same step; they drift apart by
    # rounding (a few 1e-7 per step), as each drifts from the oracle
    assert d <= RTOL_STEP and max(errs) <= RTOL_STEP
    cplx = ex.stepper.Dispersion(2, 2.0, 100, 1e-3, dispersivity=0.01)
    with pytest.raises(NotImplementedError):
        cplx.plan(1).set_rollout_strategy(2)
