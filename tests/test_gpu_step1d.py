"""GPU parity of the fused 1-D ETDRK step against the FP32 oracle.

Every (nonlinear kind, ETDRK order, N) runs a small seeded batch through the
C ABI (via the stepper classes) and is compared with the oracle's FP32
restatement of the reference step on the same inputs; rel-L2 (the
reference's nRMSE) must be <= 1e-5 after one and after ten steps.
"""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


def _pair(kind, N, order, L=None, dt=None):
    if kind == "ks_cons":
        L = L or 60.0
        dt = dt or 0.1
        return (ex.stepper.KuramotoSivashinskyConservative(1, L, N, dt, order=order),
                ost.KuramotoSivashinskyConservative(1, L, N, dt, order=order, dtype=np.float32), L)
    if kind == "ks":
        L = L or 60.0
        dt = dt or 0.05
        return (ex.stepper.KuramotoSivashinsky(1, L, N, dt, order=order),
                ost.KuramotoSivashinsky(1, L, N, dt, order=order, dtype=np.float32), L)
    if kind == "burgers":
        L = L or 1.0
        dt = dt or 1e-3
        return (ex.stepper.Burgers(1, L, N, dt, diffusivity=3e-3, order=order),
                ost.Burgers(1, L, N, dt, diffusivity=3e-3, order=order, dtype=np.float32), L)
    if kind == "burgers_exact":
        # literal u u_x with two inverse transforms (EXMI_FLAG_EXACT_FORM)
        g, c, L = _pair("burgers", N, order, L, dt)
        g.exact_form = True
        return g, c, L
    if kind == "burgers_df05":
        # a non-default dealias cutoff: the rollout kernels take it at run time
        L = L or 1.0
        dt = dt or 1e-3
        return (ex.stepper.Burgers(1, L, N, dt, diffusivity=3e-3, order=order, dealiasing_fraction=0.5),
                ost.Burgers(1, L, N, dt, diffusivity=3e-3, order=order, dealiasing_fraction=0.5,
                            dtype=np.float32), L)
    if kind == "burgers_cons":
        L = L or 1.0
        dt = dt or 1e-3
        return (ex.stepper.Burgers(1, L, N, dt, diffusivity=3e-3, order=order, conservative=True,
                                   single_channel=True),
                ost.Burgers(1, L, N, dt, diffusivity=3e-3, order=order, conservative=True, single_channel=True,
                            dtype=np.float32), L)
    raise ValueError(kind)


@pytest.mark.parametrize("N", [32, 64, 256, 1024, 2048])
@pytest.mark.parametrize("kind", ["ks_cons", "ks", "burgers", "burgers_exact", "burgers_cons"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_step_parity(device, kind, N, order):
    print(f"[case] step kind={kind} N={N} order={order}", flush=True)
    gpu, cpu, L = _pair(kind, N, order)
    u0, _ = ics(4, 1, L, N, 5, seed=11)
    u_g = ex.vmap(gpu)(u0)
    u_c = cpu.step_batch(u0)
    assert np.isfinite(u_g).all()
    assert max_rel_l2(u_g, u_c) <= RTOL_STEP
    for _ in range(9):
        u_g = ex.vmap(gpu)(u_g)
        u_c = cpu.step_batch(u_c)
    assert max_rel_l2(u_g, u_c) <= RTOL_STEP


@pytest.mark.parametrize("N", [64, 256, 1024])
def test_etdrk0_linear_exact(device, N):
    """ETDRK0 on diffusion reproduces the analytic decay (reference
    tests/test_etdrk.py:20-45) and the oracle."""
    L, nu, dt = 10.0, 0.1, 0.1

    class Diffusion(ex.BaseStepper):
        def __init__(self):
            super().__init__(1, L, N, dt, num_channels=1, order=0)

        def _build_linear_operator(self, d):
            return nu * ex.spectral.build_laplace_operator(d)

        def _build_nonlinear_fun(self, d):
            return ex.nonlin_fun.ZeroNonlinearFun(1, N)

    x = ex.make_grid(1, L, N)
    k = 4 * 2 * np.pi / L
    u0 = np.sin(k * x).astype(np.float32)
    u1 = Diffusion()(u0)
    exact = np.exp(-k * k * nu * dt) * np.sin(k * x)
    assert np.max(np.abs(u1 - exact)) < 1e-5


def test_fft_roundtrip_and_numpy(device):
    rng = np.random.default_rng(0)
    for N in [16, 64, 256, 1024, 4096]:
        print(f"[case] fft N={N}", flush=True)
        x = rng.standard_normal((3, N)).astype(np.float32)
        X = ex.fft(x)
        ref = np.fft.rfft(x.astype(np.float64))
        assert np.max(np.abs(X - ref)) <= 2e-5 * np.max(np.abs(ref)) * np.log2(N)
        y = ex.ifft(X, num_points=N)
        assert np.max(np.abs(y - x)) < 1e-5 * np.log2(N)
        # c2r ignores Im at DC and Nyquist
        X2 = ref.astype(np.complex64).copy()
        X2[:, 0] += 3j
        X2[:, -1] -= 2j
        assert np.max(np.abs(ex.ifft(X2, num_points=N) - np.fft.irfft(ref, n=N))) < 1e-5 * np.log2(N)


def test_shape_error(device):
    s = ex.stepper.Burgers(1, 1.0, 64, 0.01)
    with pytest.raises(ValueError, match="Expected shape"):
        s(np.zeros((2, 64), np.float32))


@pytest.mark.parametrize("N", [64, 256, 1024, 2048])
@pytest.mark.parametrize("kind", ["ks_cons", "ks", "burgers", "burgers_exact", "burgers_df05"])
@pytest.mark.parametrize("order", [2, 4])
def test_rollout_strategies_agree(device, kind, N, order):
    """The multi-step rollout kernels (1 or 2 transforms per lane group) and
    the per-step launches compute the same steps; all match the oracle
    rollout (reference _utils.py:92-186) at steps 1 and 10. ETDRK2 with the
    default 2/3 cutoff runs the folded constant-cutoff kernels for the three
    kinds that have them (conservative convection, the exact u u_x form,
    gradient norm); the 0.5 cutoff runs the general ones."""
    import torch

    print(f"[case] rollout kind={kind} N={N} order={order}", flush=True)
    gpu, cpu, L = _pair(kind, N, order)
    B, n = 5, 12
    u0, _ = ics(B, 1, L, N, 5, seed=21)
    x = torch.tensor(u0, device=device)
    plan = gpu.plan(B)
    outs = {}
    for strat in (-1, 1, 2):
        try:
            plan.set_rollout_strategy(strat)
        except NotImplementedError:
            continue
        outs[strat] = gpu.rollout_batch(x, n, include_init=True).cpu().numpy()
    ref = outs[-1]
    # ETDRK2 plans whose dealias mask removes the Nyquist mode keep the
    # spectrum on chip in the multi-step kernels (exact identity of the
    # per-step round trip, float32 rounding apart); the per-step launches
    # (-1) keep the round trip
    tol_rt = 5e-6 if order == 2 else 1e-6
    for strat, tr in outs.items():
        assert tr.shape == (B, n + 1, 1, N)
        np.testing.assert_array_equal(tr[:, 0], u0)
        tol = 1e-6 if strat == -1 else tol_rt
        assert max_rel_l2(tr[:, 1:].reshape(B * n, 1, N), ref[:, 1:].reshape(B * n, 1, N)) <= tol, strat
    if 1 in outs and 2 in outs:
        assert max_rel_l2(outs[1][:, 1:].reshape(B * n, 1, N), outs[2][:, 1:].reshape(B * n, 1, N)) <= 1e-6
    u = u0
    for s in range(1, 11):
        u = cpu.step_batch(u)
        if s in (1, 10):
            for tr in outs.values():
                assert max_rel_l2(tr[:, s], u) <= RTOL_STEP


def test_repeat_matches_rollout_last(device):
    import torch

    gpu, _, L = _pair("ks_cons", 256, 2)
    u0, _ = ics(3, 1, L, 256, 5, seed=5)
    x = torch.tensor(u0, device=device)
    tr = ex.rollout(ex.vmap(gpu), 7)(x)
    assert tuple(tr.shape) == (7, 3, 1, 256)  # time first, as lax.scan over the batched state
    last = ex.repeat(ex.vmap(gpu), 7)(x)
    assert max_rel_l2(last.cpu().numpy(), tr[-1].cpu().numpy()) <= 1e-6
    assert max_rel_l2(tr.movedim(0, 1).cpu().numpy().reshape(21, 1, 256),
                      gpu.rollout_batch(x, 7).cpu().numpy().reshape(21, 1, 256)) == 0.0


@pytest.mark.parametrize("kind", ["ks_cons", "burgers"])
def test_fourier_resident_rollout(device, kind):
    """The Fourier-resident rollout (EXMI_FLAG_FOURIER_RESIDENT) stays within
    the parity gate of the oracle at steps 1 and 10 and within float32
    rounding of the physical rollout; both multi-step variants agree."""
    import torch

    print(f"[case] fourier-resident {kind}", flush=True)
    N = 1024 if kind == "burgers" else 256
    gpu, cpu, L = _pair(kind, N, 2)
    gpu.fourier_resident = False  # the per-step physical round trip
    fr, _, _ = _pair(kind, N, 2)
    fr.fourier_resident = True
    u0, _ = ics(6, 1, L, N, 5, seed=13)
    x = torch.tensor(u0, device=device)
    a = gpu.rollout_batch(x, 10).cpu().numpy()
    outs = []
    for nt in (1, 2):
        fr.plan(6).set_rollout_strategy(nt)
        outs.append(fr.rollout_batch(x, 10).cpu().numpy())
    assert max_rel_l2(outs[1][:, -1], outs[0][:, -1]) <= 1e-6
    b = outs[0]
    u = u0
    for s in range(1, 11):
        u = cpu.step_batch(u)
        if s in (1, 10):
            assert max_rel_l2(b[:, s - 1], u) <= RTOL_STEP
    assert max_rel_l2(b[:, -1], a[:, -1]) <= 1e-6
    last = fr.repeat_batch(x, 10).cpu().numpy()
    assert max_rel_l2(last, b[:, -1]) <= 1e-6


@pytest.mark.parametrize("kind,N", [("ks_cons", 256), ("burgers", 1024)])
@pytest.mark.parametrize("physical", [True, False])
def test_rollout_resume_equals_continuous(device, kind, N, physical):
    """Checkpoint / resume: a rollout of 100 steps resumed from its last state
    for 100 more equals the continuous 200-step rollout. With the per-step
    physical round trip the multi-step kernel does the same arithmetic per
    step, so the match is bit for bit; the default (spectrum kept on chip)
    restarts from the rounded physical state and matches to rounding."""
    import torch

    gpu, _, L = _pair(kind, N, 2)
    if physical:
        gpu.fourier_resident = False
    u0, _ = ics(5, 1, L, N, 5, seed=14)
    x = torch.tensor(u0, device=device)
    full = gpu.rollout_batch(x, 200)
    a = gpu.rollout_batch(x, 100)
    b = gpu.rollout_batch(a[:, -1].contiguous(), 100)
    if physical:
        torch.testing.assert_close(torch.cat([a, b], dim=1), full, rtol=0, atol=0)
    else:
        got = torch.cat([a, b], dim=1).cpu().numpy()
        assert max_rel_l2(got.reshape(-1, 1, N), full.cpu().numpy().reshape(-1, 1, N)) <= 1e-5


def test_autotuner_cache_roundtrip(device, tmp_path, monkeypatch):
    """EXMI_TUNE_CACHE: a tuned strategy is written once and a new plan of
    the same shape takes it from the file without probing."""
    import torch

    path = tmp_path / "tune.txt"
    monkeypatch.setenv("EXMI_TUNE_CACHE", str(path))
    u0, _ = ics(3, 1, 60.0, 128, 5, seed=15)
    x = torch.tensor(u0, device=device)
    s1 = ex.stepper.KuramotoSivashinskyConservative(1, 60.0, 128, 0.1)
    s1.rollout_batch(x, 20)
    chosen = s1.plan(3).rollout_strategy()
    lines = path.read_text().split("\n")
    assert len([ln for ln in lines if ln.strip()]) == 1 and lines[0].endswith(f" {chosen}")
    path.write_text(lines[0].rsplit(" ", 1)[0] + " 1\n")  # force a different, valid choice
    s2 = ex.stepper.KuramotoSivashinskyConservative(1, 60.0, 128, 0.1)
    out = s2.rollout_batch(x, 20)
    assert s2.plan(3).rollout_strategy() == 1
    assert len([ln for ln in path.read_text().split("\n") if ln.strip()]) == 1  # nothing re-tuned
    assert max_rel_l2(out[:, -1].cpu().numpy(), s1.rollout_batch(x, 20)[:, -1].cpu().numpy()) <= 1e-5


def test_autotuner_short_first_rollout_is_provisional(device, tmp_path, monkeypatch):
    """A first rollout shorter than the 16-step probe chooses a strategy for
    itself only: nothing is stored, and the first long rollout times the
    candidates again and stores that choice (2-D Kolmogorov at 128^2, where
    launch counts weigh on short probes)."""
    import torch

    path = tmp_path / "tune.txt"
    monkeypatch.setenv("EXMI_TUNE_CACHE", str(path))
    st = ex.stepper.KolmogorovFlowVorticity(2, 2 * np.pi, 128, 0.01, diffusivity=1e-2)
    u0, _ = ics(4, 2, 2 * np.pi, 128, 6, seed=16)
    x = torch.tensor(u0, device=device)
    short = st.rollout_batch(x, 2)
    assert st.plan(4).rollout_strategy() != 0
    assert not path.exists() or not path.read_text().strip()
    long = st.rollout_batch(x, 20)
    lines = [ln for ln in path.read_text().split("\n") if ln.strip()]
    assert len(lines) == 1 and lines[0].endswith(f" {st.plan(4).rollout_strategy()}")
    assert max_rel_l2(long[:, 1].cpu().numpy(), short[:, 1].cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("order", [2, 3])
def test_repeated_stepper(device, order):
    """RepeatedStepper (reference _repeated_stepper.py:9-143): dt is the
    sub-step dt times num_sub_steps; a call equals the oracle's transform-once
    sub-stepping (fft, num_sub_steps x step_fourier, ifft) and the repeat of
    the plain stepper; rollouts of the wrapper return every call's state."""
    import torch

    from oracle import spectral as osp

    k = 6
    gpu, cpu, L = _pair("burgers", 256, order)
    rs = ex.RepeatedStepper(gpu, k)
    assert rs.dt == pytest.approx(gpu.dt * k, abs=1e-12)
    u0, _ = ics(3, 1, L, 256, 5, seed=31)
    x = torch.tensor(u0, device=device)
    out = rs.step(x).cpu().numpy()
    uh = osp.fft(u0, num_spatial_dims=1).astype(np.complex64)
    for _ in range(k):
        uh = cpu.step_fourier(uh)
    ref = osp.ifft(uh, num_spatial_dims=1, num_points=256).astype(np.float32)
    assert max_rel_l2(out, ref) <= RTOL_STEP
    assert max_rel_l2(out, gpu.repeat_batch(x, k).cpu().numpy()) <= 1e-6
    with pytest.raises(ValueError):
        rs(x)  # a batch through __call__
    single = rs(x[0]).cpu().numpy()
    assert max_rel_l2(single[None], out[:1]) <= 1e-6
    traj = ex.rollout(ex.vmap(rs), 3, include_init=True)(x).cpu().numpy()
    assert traj.shape == (4, 3, 1, 256)  # time first
    np.testing.assert_array_equal(traj[0], u0)
    np.testing.assert_array_equal(traj[1], out)
    last = ex.repeat(ex.vmap(rs), 3)(x).cpu().numpy()
    np.testing.assert_array_equal(last, traj[3])


def test_rollout_to_file(device, tmp_path):
    """Chunked device rollouts streamed to a .npy dataset equal one
    continuous rollout: bit for bit with the per-step physical round trip,
    to float32 rounding with the default spectrum-resident rollout (each
    chunk restarts from the stored float32 state); the sidecar records the
    stepper."""
    import torch

    gpu, _, L = _pair("ks_cons", 256, 2)
    u0, _ = ics(4, 1, L, 256, 5, seed=33)
    x = torch.tensor(u0, device=device)
    path = str(tmp_path / "ks_default.npy")
    ex.io.rollout_to_file(gpu, x, 25, path, chunk=7)
    arr, _ = ex.io.open_trajectories(path)
    full = gpu.rollout_batch(x, 25, include_init=True).cpu().numpy()
    assert max_rel_l2(np.asarray(arr).reshape(-1, 1, 256), full.reshape(-1, 1, 256)) <= 1e-6
    gpu, _, _ = _pair("ks_cons", 256, 2)
    gpu.fourier_resident = False  # before the first plan is built
    path = str(tmp_path / "ks.npy")
    ex.io.rollout_to_file(gpu, x, 25, path, chunk=7, meta={"ic": "TFS cutoff 5 seed 33"})
    arr, side = ex.io.open_trajectories(path)
    full = gpu.rollout_batch(x, 25, include_init=True).cpu().numpy()
    np.testing.assert_array_equal(np.asarray(arr), full)
    assert side["stepper"] == "KuramotoSivashinskyConservative" and side["steps"] == 25
    assert side["shape"] == [4, 26, 1, 256] and side["user"]["ic"].startswith("TFS")
    rs = ex.RepeatedStepper(gpu, 3)
    p2 = str(tmp_path / "ks_rep.npy")
    ex.io.rollout_to_file(rs, x, 4, p2, chunk=3)
    arr2, side2 = ex.io.open_trajectories(p2)
    np.testing.assert_array_equal(np.asarray(arr2), rs.rollout_batch(x, 4, include_init=True).cpu().numpy())
    assert side2["num_sub_steps"] == 3 and side2["sub_stepper"] == "KuramotoSivashinskyConservative"


def test_ks_cons_chaotic_statistics_500_steps(device):
    """The plain configuration's long rollout (KS-conservative 1-D, L = 100,
    N = 256, dt = 0.1, ETDRK2, IC TFS cutoff 5 seed 1, 500 steps), judged by
    statistics as a chaotic trajectory: the time-averaged power spectrum
    (get_spectrum, _spectral.py:866-1030) over steps 251-500 against the
    float64 oracle's over the same window, modes 1..kmax, rel-L2 <= 1e-2, and
    the mean energy 0.5 mean(u^2) over that window within 1e-3 relative.
    The float32 oracle itself sits at 4.8e-4 and 3e-5 from the float64 one
    on this window (the gates are ~20x that). The per-step distance to the
    float64 oracle is printed at steps 1, 10, 100 and 500."""
    import torch

    from oracle import spectral as osp

    L, N, DT, K = 100.0, 256, 0.1, 500
    u0, _ = ics(1, 1, L, N, 5, seed=1)
    gpu = ex.stepper.KuramotoSivashinskyConservative(1, L, N, DT)
    tr = gpu.rollout_batch(torch.tensor(u0, device=device), K)
    h0 = K // 2
    spec_gpu = ex.get_spectrum(tr[0, h0:]).mean(dim=0)[0].cpu().numpy().astype(np.float64)
    e_gpu = float(0.5 * (tr[0, h0:].double() ** 2).mean().item())
    g = tr[0].cpu().numpy().astype(np.float64)
    o64 = ost.KuramotoSivashinskyConservative(1, L, N, DT)
    u = u0[0].astype(np.float64)
    spec_ref = np.zeros(N // 2 + 1)
    e_ref = []
    for s in range(1, K + 1):
        u = o64.step_batch(u)
        if s in (1, 10, 100, 500):
            print(f"  step {s}: rel-L2 vs float64 oracle {ost.nrmse(g[s - 1], u):.2e}", flush=True)
        if s > h0:
            spec_ref += osp.get_spectrum(u)[0]
            e_ref.append(0.5 * np.mean(u * u))
    spec_ref /= K - h0
    kmax = int(2 / 3 * (N // 2)) - 1
    es = np.linalg.norm(spec_gpu[1:kmax + 1] - spec_ref[1:kmax + 1]) / np.linalg.norm(spec_ref[1:kmax + 1])
    ee = abs(e_gpu - np.mean(e_ref)) / np.mean(e_ref)
    print(f"  time-averaged spectrum rel-L2 {es:.2e}, mean energy rel {ee:.2e}", flush=True)
    assert np.isfinite(g).all()
    assert es <= 1e-2 and ee <= 1e-3


@pytest.mark.parametrize("name", ["ks_cons", "ks", "burgers", "burgers_cons", "burgers_df05", "allen_cahn",
                                  "cahn_hilliard", "general", "kdv"])
def test_nonlinear_fun_1d(device, name):
    """The nonlinear functions on 1-D spectra (BaseNonlinearFun.__call__,
    nonlin_fun/_base.py:139-159) against the oracle's: single-channel
    conservative convection, gradient norm, the non-conservative convection
    (in its literal two-transform form), a 0.5 dealias fraction, the
    polynomial, Cahn-Hilliard's Laplacian of the cubic, the general form and
    the convection of a complex-L plan (KdV); max |error| / max |N| <= 1e-5."""
    import torch

    from oracle import spectral as osp

    print(f"[case] 1-D nonlinear function {name}", flush=True)
    N = 256
    if name in ("ks_cons", "ks", "burgers", "burgers_cons", "burgers_df05"):
        gpu, cpu, L = _pair(name, N, 2)
    elif name == "allen_cahn":
        L = 1.0
        gpu, cpu = ex.stepper.reaction.AllenCahn(1, L, N, 1e-3), ost.AllenCahn(1, L, N, 1e-3, dtype=np.float32)
    elif name == "cahn_hilliard":
        L = 1.0
        gpu = ex.stepper.reaction.CahnHilliard(1, L, N, 1e-5)
        cpu = ost.CahnHilliard(1, L, N, 1e-5, dtype=np.float32)
    elif name == "general":
        L = 2.0
        gpu = ex.stepper.generic.GeneralNonlinearStepper(1, L, N, 1e-3, nonlinear_coefficients=(0.3, -1.0, -0.5))
        cpu = ost.GeneralStepper(1, L, N, 1e-3, nonlinear_scales=(0.3, -1.0, -0.5), dtype=np.float32)
    else:
        L = 20.0
        gpu, cpu = ex.stepper.KortewegDeVries(1, L, N, 1e-3), ost.KortewegDeVries(1, L, N, 1e-3, dtype=np.float32)
    u0, _ = ics(5, 1, L, N, 8, seed=88)
    uh = osp.fft(u0, num_spatial_dims=1).astype(np.complex64)
    got = gpu.nonlinear_fun(torch.tensor(uh, device=device)).cpu().numpy()
    want = cpu.nonlinear_fun(uh)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  max rel {err:.2e}", flush=True)
    assert got.shape == want.shape
    assert err <= 1e-5


def test_dlpack_producer_round_trip(device):
    """Arrays from a DLPack producer that is neither torch nor NumPy (the
    role of a JAX array): a device producer is used in place (zero copy), a
    host producer is moved to the device, and both step like the torch
    tensor they wrap."""
    import torch

    class Producer:
        """The DLPack protocol and a shape, as a JAX array has."""

        def __init__(self, t):
            self._t = t
            self.shape = tuple(t.shape)

        def __dlpack__(self, stream=None):
            return self._t.__dlpack__(stream=stream) if stream is not None else self._t.__dlpack__()

        def __dlpack_device__(self):
            return self._t.__dlpack_device__()

    gpu, _, L = _pair("ks_cons", 256, 2)
    u0, _ = ics(3, 1, L, 256, 5, seed=41)
    x = torch.tensor(u0, device=device)
    want = gpu.step(x).cpu().numpy()
    from exponax_amd._tensor import as_device_tensor

    t, info = as_device_tensor(Producer(x))
    assert info[0] == "dlpack" and t.data_ptr() == x.data_ptr()  # no copy of a device array
    got = gpu.step(Producer(x))
    got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
    np.testing.assert_array_equal(got, want)
    got_host = gpu.step(Producer(torch.tensor(u0)))
    got_host = got_host.cpu().numpy() if hasattr(got_host, "cpu") else np.asarray(got_host)
    np.testing.assert_array_equal(got_host, want)
    one = gpu(Producer(x[0]))
    one = one.cpu().numpy() if hasattr(one, "cpu") else np.asarray(one)
    np.testing.assert_array_equal(one, want[0])


@pytest.mark.parametrize("physical", [False, True])
def test_rollout_checked_flags_first_bad_step(device, physical):
    """exmi_rollout_checked: the multi-step kernel (or, with the per-step
    launches, a scan of the written states) reports per trajectory the first
    time index whose state is non-finite or leaves the bound, -1 otherwise;
    the trajectory equals the unchecked rollout's. A state forced to
    blow up from the host (a huge initial amplitude) and a bound crossed by
    another one are caught at the step the oracle's own states show it."""
    import torch

    gpu, cpu, L = _pair("burgers", 256, 2)
    gpu.fourier_resident = None if not physical else False
    u0, _ = ics(4, 1, L, 256, 5, seed=57)
    u0[1] *= 1e18  # overflows to inf within a few steps
    u0[2] *= 3.0
    x = torch.tensor(u0, device=device)
    n = 12
    plan = gpu.plan(4)
    plan.set_rollout_strategy(-1 if physical else 1)
    tr, fb = gpu.rollout_checked(x, n, bound=2.5, include_init=True)
    ref = gpu.rollout_batch(x, n, include_init=True)
    torch.testing.assert_close(tr[:, 1:], ref[:, 1:], rtol=0, atol=0, equal_nan=True)
    # expected from the written states themselves (t >= 1)
    want = []
    for b in range(4):
        s = tr[b, 1:].reshape(n, -1)
        bad = ~(s.abs() <= 2.5).all(dim=1)
        want.append(int(torch.nonzero(bad)[0].item()) + 1 if bool(bad.any()) else -1)
    print(f"  first bad {fb.tolist()} (expected {want})", flush=True)
    assert fb.tolist() == want and want[0] == -1 and want[1] >= 1 and want[2] >= 1
    # the scan on its own, non-finite values only
    fb2 = ex.metrics.first_bad_step(tr[:, 1:])
    want2 = []
    for b in range(4):
        s = tr[b, 1:].reshape(n, -1)
        bad = ~torch.isfinite(s).all(dim=1)
        want2.append(int(torch.nonzero(bad)[0].item()) if bool(bad.any()) else -1)
    assert fb2.tolist() == want2 and want2[1] >= 0 and want2[0] == -1
