"""Throughput benchmark of the fused ETDRK step on one MI355X.

Default workload (the batched Burgers rollout): Burgers 1-D, N=1024, L=1,
dt=1e-3, ETDRK2, 4096 trajectories advanced together, per-trajectory
viscosity nu log-uniform in [1e-4, 1e-2], float32 truncated-Fourier ICs
(cutoff 8, seed 2), a rollout of 1000 steps captured in one hipGraph, every
state written to the trajectory buffer (rollout semantics).

A bench step is one replay of the workload's rollout graph (1000 ETDRK steps
of all 4096 trajectories, the configuration's unit of work); `value` is
state-steps/s = trajectories * 1000 * bench steps / wall time of the timed
replays, with the ICs already resident in HBM. The rollout runs as the
library's `rollout` does by default: for this plan (real linear operator,
Nyquist mode removed by the 2/3 dealias mask) the per-step forward transform
of the emitted state is an exact identity, so the multi-step kernel keeps
each trajectory's spectrum on chip and writes its inverse transform every
step; the literal per-step physical round trip of the reference is measured
beside it (`physical_roundtrip`). The strategy is the one the library's
autotuner picks (one multi-step launch, or per-step launches). The roofline
uses the algorithmic bytes of one launch: trajectories * steps-per-launch *
(read + write of the state + 12 B per mode of per-trajectory ETDRK2 tables);
`roofline.traffic` is the HBM traffic of the same launch from the committed
rocprofv3 FETCH_SIZE/WRITE_SIZE summary (bytes per state-step, copy-probe
calibrated), scaled to this run's state-steps (`traffic_source`).

The CPU baseline (rank 0 only) runs the NumPy/SciPy FP32 restatement of the
same step on a bounded sample of the workload and doubles as the parity
check of the GPU trajectory at steps 1 and 10.

Other workloads (--workload): ks1d (Kuramoto-Sivashinsky conservative 1-D,
N=256, one trajectory, 500 steps), kolmogorov2d (Kolmogorov vorticity,
32 x 128^2 + 16 x 256^2 + 8 x 512^2 instances with per-instance Re, 2000
steps in one graph), ks2d (Kuramoto-Sivashinsky 2-D 128^2, 32 trajectories,
50 steps), ks2d_grad (the gradient of an MSE loss through those 50 steps,
checkpoints every 10), corrector (the learned corrector: GPU training data,
GPU training, corrected coarse rollout of 8192 trajectories against the
plain fine stepper), ns3d (Navier-Stokes velocity 3-D, 64^3, 8 trajectories
of three channels, 50 steps). They print the same JSON contract for their
own workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "state-steps/sec at rel-L2≤1e-5 vs ref; %HBM peak; learned-vs-plain speedup"
HBM_PEAK_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))


def tfs_ics(batch, d, n, length, cutoff, rng):
    """The truncated-Fourier-series IC recipe (identical to the test
    fixtures): modes with 1 <= max_d |k_d| <= cutoff from one half-plane,
    a ~ N(0,1), phase ~ U[0, 2pi), u = sum a cos(2 pi k.x / L + phase),
    u /= max|u|; trajectories drawn in order from one stream."""
    import itertools

    modes = [k for k in itertools.product(range(-cutoff, cutoff + 1), repeat=d)
             if max(abs(c) for c in k) >= 1 and [c for c in reversed(k) if c != 0][0] > 0]
    modes = np.array(modes, dtype=np.float64)
    g = np.linspace(0.0, length, n, endpoint=False)
    grids = np.meshgrid(*([g] * d), indexing="ij")
    out = np.empty((batch, 1) + (n,) * d)
    for b in range(batch):
        a = rng.standard_normal(len(modes))
        ph = rng.uniform(0.0, 2 * np.pi, len(modes))
        u = np.zeros((n,) * d)
        for m in range(len(modes)):
            arg = sum(modes[m, i] * grids[i] for i in range(d))
            u += a[m] * np.cos(2 * np.pi * arg / length + ph[m])
        out[b, 0] = u / np.max(np.abs(u))
    return out


def read_traffic(workload, state_steps):
    """HBM bytes of this run's launch from the committed rocprofv3 --pmc
    summary (bytes per state-step times state-steps), and where they come
    from."""
    name = f"traffic_{workload}.json"
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        per_step = json.load(f).get("bytes_per_state_step")
    if per_step is None:
        return None, None
    return per_step * state_steps, (f"profiles/{name}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same "
                                    f"launch ({per_step:.0f} B per state-step, copy-probe calibrated), times this "
                                    "run's state-steps")


def rel_l2_rows(g, ref):
    """Per-trajectory sum over channels of ||g - ref|| / ||ref||
    (metrics/_spatial.py:537-586), finite trajectories only."""
    g = np.asarray(g, dtype=np.float64).reshape(g.shape[0], -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(ref.shape[0], -1)
    ok = np.isfinite(g).all(axis=1) & np.isfinite(ref).all(axis=1)
    return np.sqrt(((g[ok] - ref[ok]) ** 2).sum(1)) / np.sqrt((ref[ok] ** 2).sum(1))


def timed_rollout(run, sync):
    """Time one call of `run` (captured graph replay) with events on the
    launching stream and host wall time, synchronized on both sides."""
    import torch

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    sync()
    t0 = time.perf_counter()
    ev0.record()
    run()
    ev1.record()
    sync()
    return time.perf_counter() - t0, ev0.elapsed_time(ev1) / 1e3


def oracle_baseline(make_oracle, u0, gpu_checks, sample_steps, label, fp64_maker=None):
    """Time the FP32 oracle on a bounded sample; compare the GPU states at
    the recorded steps; optionally the FP64 oracle error of the GPU."""
    from oracle import spectral as osp

    cores = os.cpu_count() or 1
    osp.set_fft_workers(cores)
    ens = make_oracle(np.float32)
    u = u0.astype(np.float32)
    ref = {}
    t0 = time.perf_counter()
    with np.errstate(all="ignore"):  # blown-up trajectories stay NaN, as in the reference
        for s in range(1, sample_steps + 1):
            u = ens.step_batch(u)
            if s in gpu_checks:
                ref[s] = u.copy()
    dt = time.perf_counter() - t0
    parity = {}
    for s, g in gpu_checks.items():
        if s in ref:
            r = rel_l2_rows(g, ref[s])
            parity[f"rel_l2_step{s}_max"] = float(r.max()) if r.size else None
            parity[f"rel_l2_step{s}_mean"] = float(r.mean()) if r.size else None
    parity["checked_trajectories"] = int(u0.shape[0])
    if fp64_maker is not None:
        e64 = fp64_maker(np.float64)
        u = u0.astype(np.float64)
        with np.errstate(all="ignore"):
            for s in range(1, max(gpu_checks) + 1):
                u = e64.step_batch(u)
                if s in gpu_checks:
                    r = rel_l2_rows(gpu_checks[s], u)
                    parity[f"fp64_rel_l2_step{s}_max"] = float(r.max()) if r.size else None
    base = {"value": u0.shape[0] * sample_steps / dt, "unit": "state-steps/s", "cores": cores, "kind": "port",
            "sample": f"NumPy/SciPy FP32 restatement, {label} ({dt:.1f} s), scipy.fft workers={cores}"}
    return base, parity


def result_line(workload, config, value, K, W, wall, per_launch_s, alg_bytes, n_launches, kernel, traffic):
    achieved = alg_bytes / per_launch_s / 1e9
    traffic, traffic_source = traffic if isinstance(traffic, tuple) else (traffic, None)
    res = {
        "metric": METRIC, "value": value, "unit": "state-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded truncated-Fourier ICs, seeded coefficients)",
        "config": dict(workload=workload, **config),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_us": per_launch_s * 1e6, "launches": n_launches, "kernel": kernel},
    }
    if traffic_source:
        res["roofline"]["traffic_source"] = traffic_source
    return res


# ------------------------------------------------------------------ workloads
def timed_replays(run, K, sync):
    """Time K back-to-back calls of `run` (graph replays) with events on the
    launching stream and host wall time, synchronized on both sides."""
    import torch

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    sync()
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(K):
        run()
    ev1.record()
    sync()
    return time.perf_counter() - t0, ev0.elapsed_time(ev1) / 1e3


def wl_burgers1d(args, ex, _lib, torch):
    """One bench step is one replay of the workload's rollout graph: 4096
    trajectories x R ETDRK steps (R = 1000, the configuration's graph length;
    --rollout-steps changes it), every state written to the trajectory
    buffer. value = 4096 * R * K / wall time of K replays."""
    N, L, DT = 1024, 1.0, 1e-3
    B, K, W = args.batch or 4096, args.steps or 20, args.warmup
    R = args.rollout_steps or 1000
    rng = np.random.default_rng(2)
    u0 = tfs_ics(B, 1, N, L, 8, rng)
    nus = 10.0 ** rng.uniform(-4.0, -2.0, B)
    lib = _lib.lib()
    dev = torch.device("cuda:0")
    u0_d = torch.tensor(u0, dtype=torch.float32, device=dev)
    stream = _lib.stream_ptr()
    per_state = 4 * N + 4 * N + 12 * (N // 2 + 1)

    def measure(physical):
        members = [ex.stepper.Burgers(1, L, N, DT, diffusivity=float(nu)) for nu in nus]
        if physical:  # the literal per-step round trip (EXMI_FLAG_PHYSICAL_ROLLOUT)
            for m in members:
                m.fourier_resident = False
        ens = ex.Ensemble(members)
        plan = ens.plan(B)
        traj = torch.empty((B, R, 1, N), dtype=torch.float32, device=dev)
        # the autotuner's probe, capture and instantiation of the graph, then W untimed replays
        _lib.check(lib.exmi_rollout_prepare(plan.handle, u0_d.data_ptr(), N, traj.data_ptr(), R, 0))
        replay = lambda: _lib.check(lib.exmi_rollout(plan.handle, u0_d.data_ptr(), N, traj.data_ptr(), R, 0, 1,  # noqa: E731
                                                     stream))
        for _ in range(W):
            replay()
        torch.cuda.synchronize()
        wall, gpu_s = timed_replays(replay, K, torch.cuda.synchronize)
        strategy = plan.rollout_strategy()
        return traj, wall, gpu_s, strategy

    traj, wall, gpu_s, strategy = measure(False)
    n_launches = R if strategy < 0 else 1  # per replay
    spl = R // n_launches
    kernel = ("step1d_kernel<512,8,NL_CONV_CONS,2>" if strategy < 0 else
              f"roll1d_kernel<512,8,NL_CONV_CONS,2,{strategy},FR,KC=340> ({spl} steps per launch)")
    res = result_line("burgers1d_n1024_b4096_etdrk2_graph",
                      {"equation": "Burgers", "num_points": N, "domain_extent": L, "dt": DT, "order": 2,
                       "batch": B, "rollout_steps_per_graph": R, "bench_step": f"one replay of the {R}-step graph",
                       "per_trajectory_viscosity": "log-uniform [1e-4, 1e-2]", "ic": "TFS cutoff 8 seed 2",
                       "hipgraph": True},
                      B * R * K / wall, K, W, wall, gpu_s / (K * n_launches), B * spl * per_state, n_launches, kernel,
                      read_traffic("burgers1d_n1024_b4096", B * spl))
    res["roofline"]["alg_bytes_per_state_step"] = per_state
    res["us_per_etdrk_step"] = wall * 1e6 / (K * R)
    res["rollout_strategy"] = strategy
    final = traj[:, R - 1]
    res["nan_trajectories"] = {"count": int((~torch.isfinite(final).all(dim=-1).all(dim=-1)).sum().item()),
                               "at_step": R, "note": "low-viscosity trajectories blow up, as the FP64 oracle's"}
    checks = {s: traj[:, s - 1].cpu().numpy() for s in (1, 10) if s <= R}
    if not args.no_physical_roundtrip:
        traj_p, wall_p, gpu_p, sp = measure(True)
        nlp = R if sp < 0 else 1
        res["physical_roundtrip"] = {
            "value": B * R * K / wall_p, "us_per_etdrk_step": wall_p * 1e6 / (K * R),
            "roofline_frac": B * (R // nlp) * per_state / (gpu_p / (K * nlp)) / 1e9 / HBM_PEAK_GBS,
            "strategy": sp,
            "kernel": ("step1d_kernel<512,8,NL_CONV_CONS,2>" if sp < 0 else
                       f"roll1d_kernel<512,8,NL_CONV_CONS,2,{sp},physical,KC=340>"),
            "max_rel_l2_vs_default_path": {s_: float(rel_l2_rows(checks[s_], traj_p[:, s_ - 1].cpu().numpy()).max())
                                           for s_ in checks}}
        del traj_p
    del traj
    if not args.no_corrector_timing:
        res["learned_vs_plain_fine_speedup"] = corrector_inference_timing(ex, torch)
    if not args.no_cpu_baseline:
        from oracle import stepper as ost

        nb, ns = min(args.cpu_sample_batch or 4096, B), args.cpu_sample_steps or 300
        mk = lambda dt: ost.Ensemble([ost.Burgers(1, L, N, DT, diffusivity=float(nu), dtype=dt)  # noqa: E731
                                      for nu in nus[:nb]], dtype=dt)
        res["cpu_baseline"], res["parity"] = oracle_baseline(mk, u0[:nb], {s_: g[:nb] for s_, g in checks.items()},
                                                              ns, f"first {nb} trajectories x {ns} steps")
    return res


def wl_ks1d(args, ex, _lib, torch):
    N, L, DT = 256, 100.0, 0.1
    K, W = args.steps or 500, args.warmup
    u0 = tfs_ics(1, 1, N, L, 5, np.random.default_rng(1))
    st = ex.stepper.KuramotoSivashinskyConservative(1, L, N, DT)
    x = torch.tensor(u0, dtype=torch.float32, device="cuda:0")
    # untimed: the autotuner's probe (a rollout of at least 16 steps, so its
    # choice is final) and the warm-up
    st.rollout_batch(x, max(W, 16))
    torch.cuda.synchronize()
    out = {}
    wall, gpu_s = timed_rollout(lambda: out.__setitem__("tr", st.rollout_batch(x, K)), torch.cuda.synchronize)
    tr = out["tr"]
    strategy = st.plan(1).rollout_strategy()
    n_launches = K if strategy < 0 else 1
    per_state = 8 * N  # operators shared: the state read and write
    res = result_line("ks_cons1d_n256_b1_etdrk2",
                      {"equation": "KuramotoSivashinskyConservative", "num_points": N, "domain_extent": L,
                       "dt": DT, "order": 2, "batch": 1, "rollout_steps": K, "ic": "TFS cutoff 5 seed 1"},
                      K / wall, K, W, wall, gpu_s / n_launches, (K // n_launches) * per_state, n_launches,
                      "roll1d_kernel<128,8,NL_CONV_CONS,2,*>" if strategy > 0 else "step1d_kernel<128,8,...>",
                      None)
    res["rollout_strategy"] = strategy
    if not args.no_cpu_baseline:
        from oracle import stepper as ost

        g = tr.cpu().numpy()
        checks = {s: g[:, s - 1] for s in (1, 10, 100, K)}
        mk = lambda dt: ost.KuramotoSivashinskyConservative(1, L, N, DT, dtype=dt)  # noqa: E731
        res["cpu_baseline"], res["parity"] = oracle_baseline(mk, u0, checks, K, f"1 trajectory x {K} steps",
                                                              fp64_maker=mk)
        # chaotic statistics over the second half of the rollout: the
        # time-averaged power spectrum (GPU spectrum kernel on the GPU states,
        # oracle spectrum on the float64 oracle's states) and the mean energy
        from oracle import spectral as osp

        h0 = K // 2
        spec_gpu = ex.get_spectrum(tr[0, h0:]).mean(dim=0)[0].cpu().numpy().astype(np.float64)
        o64 = ost.KuramotoSivashinskyConservative(1, L, N, DT)
        u = u0[0].astype(np.float64)
        spec_ref = np.zeros(N // 2 + 1)
        e_ref = []
        for s_ in range(1, K + 1):
            u = o64.step_batch(u)
            if s_ > h0:
                spec_ref += osp.get_spectrum(u)[0]
                e_ref.append(0.5 * np.mean(u * u))
        spec_ref /= K - h0
        kmax = int(2 / 3 * (N // 2)) - 1
        e_gpu = float(0.5 * (tr[0, h0:] ** 2).mean().item())
        res["chaotic_statistics"] = {
            "window_steps": [h0 + 1, K],
            "spectrum_rel_l2_modes_1_to_kmax": float(np.linalg.norm(spec_gpu[1:kmax + 1] - spec_ref[1:kmax + 1])
                                                      / np.linalg.norm(spec_ref[1:kmax + 1])),
            "mean_energy_gpu": e_gpu, "mean_energy_fp64_oracle": float(np.mean(e_ref)),
        }
    return res


def wl_kolmogorov2d(args, ex, _lib, torch):
    K, W = args.steps or 2000, args.warmup
    groups = [(128, 32), (256, 16), (512, 8)]
    rng = np.random.default_rng(3)
    steppers, u0s, res_list = [], [], []
    for n, b in groups:
        re = rng.uniform(100.0, 1000.0, b)
        u0 = tfs_ics(b, 2, n, 2 * np.pi, 6, rng)
        steppers.append(ex.ensemble(lambda r, n=n: ex.stepper.KolmogorovFlowVorticity(2, 2 * np.pi, n, 0.01,
                                                                                      diffusivity=1.0 / r), re))
        u0s.append(torch.tensor(u0, dtype=torch.float32, device="cuda:0"))
        res_list.append((n, re, u0))
    if W > 0:
        ex.rollout_group(steppers, u0s, W)
    torch.cuda.synchronize()
    # one trajectory buffer per grid size, reused by every launch mode
    trs = [torch.empty((b, K, 1, n, n), dtype=torch.float32, device="cuda:0") for n, b in groups]
    modes = {1: "one graph, a branch per grid size", 2: "a graph per grid size, concurrent streams",
             3: "direct launches, concurrent streams"}
    walls = {}
    for mode in modes:
        ex.rollout_group(steppers, u0s, K, use_graph=mode, out=trs)  # capture / warm
        walls[mode] = timed_rollout(lambda: ex.rollout_group(steppers, u0s, K, use_graph=mode, out=trs),
                                    torch.cuda.synchronize)
    best = min(walls, key=lambda m: walls[m][0])
    wall, gpu_s = walls[best]
    nst = sum(b for _, b in groups)
    per_step_bytes = sum(b * (8 * n * n + 12 * n * (n // 2 + 1)) for n, b in groups)
    res = result_line("kolmogorov2d_mixed_32x128_16x256_8x512_etdrk2",
                      {"equation": "KolmogorovFlowVorticity", "instances": "32x128^2 + 16x256^2 + 8x512^2",
                       "domain_extent": "2 pi", "dt": 0.01, "order": 2, "reynolds": "U[100, 1000]",
                       "rollout_steps": K, "ic": "TFS cutoff 6 seed 3"},
                      nst * K / wall, K, W, wall, gpu_s / K, per_step_bytes, K,
                      "row2d/col2d kernels per grid size: the spectrum-carrying chain (4 launches per step), "
                      "the workgroup-resident kernel or the six-launch chain, as the autotuner picks",
                      read_traffic("kolmogorov2d_mixed", nst))
    # a member tuned to the cooperative launch (3) runs the launch chain (2)
    # inside a group of several plans
    res["rollout_strategy_by_grid"] = {f"{n}^2": (lambda s: 2 if s == 3 else s)(st.plan(b).rollout_strategy())
                                       for st, (n, b) in zip(steppers, groups)}
    res["nan_trajectories"] = int(sum((~torch.isfinite(t[:, -1]).reshape(t.shape[0], -1).all(1)).sum().item()
                                      for t in trs))
    res["launch_mode"] = modes[best]
    res["ms_per_step_by_launch_mode"] = {modes[m]: w[0] * 1e3 / K for m, w in walls.items()}
    if not args.no_cpu_baseline:
        from oracle import stepper as ost

        par, base = {}, None
        for (n, re, u0), tr in zip(res_list, trs):
            nb = 2
            g = tr[:nb].cpu().numpy()
            checks = {s: g[:, s - 1] for s in (1, 10, 100) if s <= K}
            mk = lambda dt, n=n, re=re: ost.Ensemble(  # noqa: E731
                [ost.KolmogorovFlowVorticity(2, 2 * np.pi, n, 0.01, diffusivity=1.0 / r, dtype=dt)
                 for r in re[:nb]], dtype=dt)
            b_, p_ = oracle_baseline(mk, u0[:nb], checks, 100, f"{nb} x {n}^2 instances x 100 steps", fp64_maker=mk)
            par[f"n{n}"] = p_
            if n == 128:
                base = b_
        res["cpu_baseline"], res["parity"] = base, par
    return res


def wl_ks2d(args, ex, _lib, torch):
    N, L, DT = 128, 60.0, 0.05
    B, K, W = args.batch or 32, args.steps or 50, args.warmup
    u0 = tfs_ics(B, 2, N, L, 5, np.random.default_rng(4))
    st = ex.stepper.KuramotoSivashinsky(2, L, N, DT)
    x = torch.tensor(u0, dtype=torch.float32, device="cuda:0")
    lib = _lib.lib()
    plan = st.plan(B)
    if args.strategy:
        plan.set_rollout_strategy(args.strategy)
    tr = torch.empty((B, K, 1, N, N), dtype=torch.float32, device="cuda:0")
    sz = N * N
    # autotuner probe, graph capture; then W untimed replays of the K-step graph
    _lib.check(lib.exmi_rollout_prepare(plan.handle, x.data_ptr(), sz, tr.data_ptr(), K, 0))
    replay = lambda: _lib.check(lib.exmi_rollout(plan.handle, x.data_ptr(), sz, tr.data_ptr(), K, 0, 1,  # noqa: E731
                                                 _lib.stream_ptr()))
    for _ in range(W):
        replay()
    torch.cuda.synchronize()
    wall, gpu_s = timed_rollout(replay, torch.cuda.synchronize)
    strategy = plan.rollout_strategy()
    kernel = {-1: "row2d/col2d kernels, per-step physical round trip (6 launches per step)",
              1: "roll2d_kernel<128, gradient norm> (workgroup-resident, one launch)",
              2: "row2d/col2d kernels, spectrum kept in the column layout (4 launches per step)",
              3: "coop2d_kernel<128, gradient norm> (the 4-pass chain as one cooperative launch)"}[strategy]
    n_launches = 1 if strategy in (1, 3) else K
    res = result_line("ks2d_n128_b32_etdrk2",
                      {"equation": "KuramotoSivashinsky", "num_points": N, "domain_extent": L, "dt": DT,
                       "order": 2, "batch": B, "rollout_steps": K, "ic": "TFS cutoff 5 seed 4"},
                      B * K / wall, K, W, wall, gpu_s / n_launches, B * 8 * N * N * (K // n_launches), n_launches,
                      kernel, None)
    res["rollout_strategy"] = strategy
    if not args.no_cpu_baseline:
        from oracle import stepper as ost

        nb = 4
        g = tr[:nb].cpu().numpy()
        checks = {s: g[:, s - 1] for s in (1, 10, K)}
        mk = lambda dt: ost.KuramotoSivashinsky(2, L, N, DT, dtype=dt)  # noqa: E731
        res["cpu_baseline"], res["parity"] = oracle_baseline(mk, u0[:nb], checks, K,
                                                              f"{nb} trajectories x {K} steps", fp64_maker=mk)
    return res


def wl_ns3d(args, ex, _lib, torch):
    """3-D Navier-Stokes in velocity form (Leray-projected rotational
    convection), N = 64, L = 2 pi, three channels, ETDRK2."""
    N, L, DT = 64, 2 * np.pi, 0.01
    B, K, W = args.batch or 8, args.steps or 50, args.warmup
    # the seeded recipe built on the device (three channels per trajectory)
    x = ex.ic.truncated_fourier_series_device(B, 3, L, N, 3, 7, num_channels=3, device="cuda:0")
    u0 = x.cpu().numpy()
    st = ex.stepper.NavierStokesVelocity(3, L, N, DT, diffusivity=0.01)
    # untimed: the autotuner's probe (a rollout of at least 16 steps, so its
    # choice is final) and the warm-up
    st.rollout_batch(x, max(W, 16))
    torch.cuda.synchronize()
    out = {}
    wall, gpu_s = timed_rollout(lambda: out.__setitem__("tr", st.rollout_batch(x, K)), torch.cuda.synchronize)
    tr = out["tr"]
    res = result_line("navier_stokes_velocity3d_n64_b8_etdrk2",
                      {"equation": "NavierStokesVelocity", "num_points": N, "domain_extent": L, "dt": DT,
                       "order": 2, "batch": B, "rollout_steps": K, "ic": "TFS cutoff 3 seed 7 (device)"},
                      B * K / wall, K, W, wall, gpu_s / K, B * 8 * 3 * N ** 3, K,
                      "3-D transform, field, product, combine and stage passes (per step)", None)
    res["rollout_strategy"] = st.plan(B).rollout_strategy()
    if not args.no_cpu_baseline:
        from oracle import stepper as ost

        nb = 1
        g = tr[:nb].cpu().numpy()
        checks = {s: g[:, s - 1] for s in (1, 10)}
        mk = lambda dt: ost.NavierStokesVelocity(3, L, N, DT, diffusivity=0.01, dtype=dt)  # noqa: E731
        res["cpu_baseline"], res["parity"] = oracle_baseline(mk, u0[:nb], checks, 10, f"{nb} trajectory x 10 steps")
    return res


def wl_kolmogorov2d_513(args, ex, _lib, torch):
    """A grid size outside the fused kernels' set: the 2-D Kolmogorov flow at
    the reference validation notebook's N = 513 (= 27 x 19), L = 2 pi,
    dt = 0.01, ETDRK2, 8 instances with Re in [100, 1000]: the multi-pass
    plan on the runtime-size mixed-radix transforms (radix/tile chosen by the
    plan's autotuner). value = instance-steps/s of a K-step rollout."""
    N, L, DT = 513, 2 * np.pi, 0.01
    B, K, W = args.batch or 8, args.steps or 50, args.warmup
    rng = np.random.default_rng(6)
    re = rng.uniform(100.0, 1000.0, B)
    u0 = tfs_ics(B, 2, N, L, 6, rng)
    st = ex.ensemble(lambda r: ex.stepper.KolmogorovFlowVorticity(2, L, N, DT, diffusivity=1.0 / r), re)
    x = torch.tensor(u0, dtype=torch.float32, device="cuda:0")
    # untimed: the autotuner's probe (a rollout of at least 16 steps, so its
    # choice is final) and the warm-up
    st.rollout_batch(x, max(W, 16))
    torch.cuda.synchronize()
    out = {}
    wall, gpu_s = timed_rollout(lambda: out.__setitem__("tr", st.rollout_batch(x, K)), torch.cuda.synchronize)
    tr = out["tr"]
    nr = N * (N // 2 + 1)
    res = result_line("kolmogorov2d_n513_b8_etdrk2",
                      {"equation": "KolmogorovFlowVorticity", "num_points": N, "domain_extent": "2 pi", "dt": DT,
                       "order": 2, "batch": B, "rollout_steps": K, "reynolds": "U[100, 1000]",
                       "ic": "TFS cutoff 6 seed 6"},
                      B * K / wall, K, W, wall, gpu_s / K, B * (8 * N * N + 12 * nr), K,
                      "runtime-size transforms (gfft_kernel) and the field, product, combine and stage passes", None)
    res["fft_config"] = st.fft_config(batch=B)
    res["rollout_strategy"] = st.plan(B).rollout_strategy()
    if not args.no_cpu_baseline:
        from oracle import stepper as ost

        nb = 1
        g = tr[:nb].cpu().numpy()
        checks = {s: g[:, s - 1] for s in (1, 10)}
        mk = lambda dt: ost.Ensemble([ost.KolmogorovFlowVorticity(2, L, N, DT, diffusivity=1.0 / re[0],  # noqa: E731
                                                                   dtype=dt)], dtype=dt)
        res["cpu_baseline"], res["parity"] = oracle_baseline(mk, u0[:nb], checks, 10, f"{nb} instance x 10 steps")
    return res


def wl_ks2d_grad(args, ex, _lib, torch):
    """Gradient of MSE(u_50, target) w.r.t. u_0 through 50 KS 2-D steps
    (128^2, 32 trajectories), checkpoints every 10 steps: the forward pass
    that keeps checkpoints, the segment recomputation and the adjoint steps,
    all on the GPU. value = trajectories * steps / wall time of one
    loss-and-gradient evaluation."""
    N, L, DT = 128, 60.0, 0.05
    B, K, W, CK = args.batch or 32, args.steps or 50, args.warmup, 10
    u0 = tfs_ics(B, 2, N, L, 5, np.random.default_rng(4))
    tg = tfs_ics(B, 2, N, L, 5, np.random.default_rng(40))
    st = ex.stepper.KuramotoSivashinsky(2, L, N, DT)
    x = torch.tensor(u0, dtype=torch.float32, device="cuda:0")
    target = torch.tensor(tg, dtype=torch.float32, device="cuda:0")

    def loss_and_grad():
        xr = x.clone().requires_grad_(True)
        y = ex.autograd.repeat(st, xr, K, checkpoint_every=CK)
        loss = ((y - target) ** 2).mean()
        loss.backward()
        return loss, xr.grad

    for _ in range(max(1, W // 10)):
        loss_and_grad()
        st.repeat_batch(x, K)
    torch.cuda.synchronize()
    out = {}
    wall_f, _ = timed_rollout(lambda: out.__setitem__("f", st.repeat_batch(x, K)), torch.cuda.synchronize)
    wall, gpu_s = timed_rollout(lambda: out.__setitem__("g", loss_and_grad()), torch.cuda.synchronize)
    loss, grad = out["g"]
    # per state-step: checkpoint pass and recomputation (8 N^2 each) plus the
    # adjoint step (reads u_t and the cotangent, writes the new cotangent)
    per_state = 8 * N * N * 2 + 12 * N * N
    coop = st.plan(B).rollout_strategy() in (0, 3)  # untuned plans run the cooperative launches too
    nseg = (K + CK - 1) // CK
    # cooperative: checkpoint pass (nseg - 1 launches), per segment one
    # recompute and one reverse-pass launch; otherwise six launches per
    # forward step and twelve per transposed step
    n_launches = (nseg - 1) + 2 * nseg if coop else K * 12 + 2 * K * 6
    kernel = ("coop2d_kernel<128, gradient norm> (checkpoint pass; segment recompute keeping the gradient "
              "fields) + coop_adj2d_kernel<128> (a segment's transposed steps, one launch)") if coop else \
        "row2d/col2d (+ adjoint modes) kernels, 12 launches per adjoint step"
    res = result_line("ks2d_n128_b32_grad50_ckpt10",
                      {"equation": "KuramotoSivashinsky", "num_points": N, "domain_extent": L, "dt": DT,
                       "order": 2, "batch": B, "rollout_steps": K, "checkpoint_every": CK,
                       "loss": "mean((u_50 - target)^2)", "ic": "TFS cutoff 5 seed 4, target seed 40"},
                      B * K / wall, K, W, wall, gpu_s / n_launches, B * per_state * K / n_launches, n_launches,
                      kernel, None)
    res["unit_note"] = "state-steps/s of a full loss+gradient evaluation"
    res["forward_only_state_steps_per_s"] = B * K / wall_f
    res["gradient_cost_over_forward"] = wall / wall_f
    res["loss"] = float(loss.item())
    if not args.no_cpu_baseline:
        from oracle import adjoint as oad
        from oracle import stepper as ost

        g = grad.cpu().numpy().astype(np.float64)
        cpu = ost.KuramotoSivashinsky(2, L, N, DT)
        par = {}
        t0 = time.perf_counter()
        nb = 2
        for b in range(nb):
            _, gref = oad.mse_grad(cpu, u0[b], tg[b], K)
            gref = gref / B
            par[f"grad_rel_l2_fp64_traj{b}"] = float(np.linalg.norm(g[b] - gref) / np.linalg.norm(gref))
        dt = time.perf_counter() - t0

        def loss64(u):
            for _ in range(K):
                u = cpu.step_batch(u)
            return float(np.mean((u - tg[:1]) ** 2))

        v = np.random.default_rng(41).standard_normal(u0[:1].shape)
        eps = 1e-3
        fd = (loss64(u0[:1] + eps * v) - loss64(u0[:1] - eps * v)) / (2 * eps) / B
        an = float(np.sum(g[:1] * v))
        par["fd_directional_rel_err"] = abs(an - fd) / abs(fd)
        res["parity"] = par
        res["cpu_baseline"] = {"value": nb * K / dt, "unit": "state-steps/s (loss+gradient)",
                               "cores": os.cpu_count() or 1, "kind": "port",
                               "sample": f"NumPy float64 reverse pass (oracle/adjoint.py), {nb} trajectories x "
                                         f"{K} steps ({dt:.1f} s)"}
    return res


CNN_FLOP_PER_SAMPLE = 2 * (1 * 32 * 5 + 3 * 32 * 32 * 5 + 32 * 1 * 5) * 64  # 2,007,040 on 64 points
BF16_PEAK_TFLOPS = 2500.0


def corrector_inference_timing(ex, torch):
    """The learned-corrector part of the metric, timed: the corrected coarse
    stepper (KS-conservative 1-D, L = 100, N = 64, dt = 0.5, 200 steps, one
    fused step and one network launch per step) against the plain fine
    stepper (N = 256, dt = 0.1, 1000 steps) over the same 100 time units on
    8192 trajectories, and the network kernel's fraction of the BF16 MFMA
    peak. The network has the trained one's shape with seeded weights: the
    wall times do not depend on the weights; the accuracy at equal rel-L2 of
    the trained network is measured by --workload corrector."""
    from exponax_amd import corrector as cr

    L, NF, NC, DTF, DTC, ST, B, K = 100.0, 256, 64, 0.1, 0.5, 5, 8192, 200
    sync = torch.cuda.synchronize
    KS = ex.stepper.KuramotoSivashinskyConservative
    fine, coarse = KS(1, L, NF, DTF), KS(1, L, NC, DTC)
    u0 = torch.tensor(ex.ic.truncated_fourier_series(B, 1, L, NF, 5, 50), dtype=torch.float32, device="cuda:0")
    u0c = ex.map_between_resolutions(u0, NC)
    cnn = cr.PeriodicCNN(seed=0)
    corrected = cr.CorrectedStepper(coarse, cnn)
    corrected.rollout_batch(u0c, 2)
    fine.rollout_batch(u0, 10)
    keep = {}
    wall_c, _ = timed_rollout(lambda: keep.__setitem__("c", corrected.rollout_batch(u0c, K)), sync)
    wall_f, _ = timed_rollout(lambda: keep.__setitem__("f", fine.rollout_batch(u0, K * ST)), sync)
    x = keep["c"][:, -1].contiguous()
    del keep
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    sync()
    ev0.record()
    for _ in range(reps):
        cr.cnn_forward(cnn.params.detach(), x)
    ev1.record()
    sync()
    cnn_s = ev0.elapsed_time(ev1) / 1e3 / reps
    tflops = B * CNN_FLOP_PER_SAMPLE / cnn_s / 1e12
    return {"value": wall_f / wall_c, "corrected_coarse_wall_s": wall_c, "plain_fine_wall_s": wall_f,
            "trajectories": B, "coarse_steps": K, "fine_steps": K * ST,
            "cnn_mfma_frac": tflops / BF16_PEAK_TFLOPS, "cnn_tflops": tflops, "cnn_us_per_launch": cnn_s * 1e6,
            "note": "wall-time ratio of the inference rollouts (seeded weights; the trained network's "
                    "accuracy is measured by --workload corrector)"}


def wl_corrector(args, ex, _lib, torch):
    """Learned corrector: KS-conservative 1-D, L = 100. The fine stepper
    (N = 256, dt = 0.1) generates 2048 training trajectories of 150 time
    units, downsampled to N = 64 every 5th step. The network is trained on
    the GPU (4-step unrolled loss through the coarse solver's adjoint). The
    corrected coarse stepper (N = 64, dt = 0.5) then rolls 8192 held-out
    trajectories 200 steps (100 time units), against the plain fine stepper
    over the same time (1000 steps). value = state-steps/s of the corrected
    rollout. Errors are measured against the float64 fine oracle at t = 5, 10
    and 20."""
    from exponax_amd import corrector as cr

    L, NF, NC, DTF, DTC, ST = 100.0, 256, 64, 0.1, 0.5, 5
    BT, B, K = 2048, args.batch or 8192, args.steps or 200
    dev = torch.device("cuda:0")
    sync = torch.cuda.synchronize
    KS = ex.stepper.KuramotoSivashinskyConservative
    fine, coarse = KS(1, L, NF, DTF), KS(1, L, NC, DTC)
    t0 = time.perf_counter()
    u0_tr = torch.tensor(ex.ic.truncated_fourier_series(BT, 1, L, NF, 5, 5), dtype=torch.float32, device=dev)
    data = cr.downsampled_trajectories(fine, u0_tr, 300, ST, NC)
    sync()
    t_data = time.perf_counter() - t0
    cnn = cr.PeriodicCNN(seed=0)
    t0 = time.perf_counter()
    hist = cr.train(cnn, coarse, data, iters=args.train_iters or 3000, batch=512, unroll=4, lr=1e-3,
                    log_every=500)
    sync()
    t_train = time.perf_counter() - t0
    del data

    u0 = torch.tensor(ex.ic.truncated_fourier_series(B, 1, L, NF, 5, 50), dtype=torch.float32, device=dev)
    u0c = ex.map_between_resolutions(u0, NC)
    corrected = cr.CorrectedStepper(coarse, cnn)
    corrected.rollout_batch(u0c, 2)
    fine.rollout_batch(u0, 10)
    coarse.rollout_batch(u0c, 2)
    out = {}
    wall_c, _ = timed_rollout(lambda: out.__setitem__("c", corrected.rollout_batch(u0c, K)), sync)
    wall_f, _ = timed_rollout(lambda: out.__setitem__("f", fine.rollout_batch(u0, K * ST)), sync)
    wall_p, _ = timed_rollout(lambda: out.__setitem__("p", coarse.rollout_batch(u0c, K)), sync)
    # the network kernel alone, events on the launching (current) stream
    x = out["c"][:, -1].contiguous()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    sync()
    ev0.record()
    for _ in range(reps):
        cr.cnn_forward(cnn.params.detach(), x)
    ev1.record()
    sync()
    cnn_s = ev0.elapsed_time(ev1) / 1e3 / reps
    achieved = B * CNN_FLOP_PER_SAMPLE / cnn_s / 1e12
    res = {
        "metric": METRIC, "value": B * K / wall_c, "unit": "state-steps/s", "n_gpus": 1, "steps": K, "warmup": 1,
        "ms_per_step": wall_c * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 solver, bf16 network (f32 accumulate)",
        "data": "synthetic (seeded truncated-Fourier ICs); network trained in this run on GPU-generated data",
        "config": {"workload": "corrector_ks_cons1d_l100_fine256_dt01_coarse64_dt05_b8192_200steps",
                   "equation": "KuramotoSivashinskyConservative", "domain_extent": L, "fine": [NF, DTF],
                   "coarse": [NC, DTC], "train_trajectories": BT, "train_horizon_coarse_steps": 300,
                   "eval_trajectories": B, "eval_coarse_steps": K, "network": "5x Conv1d k5 32ch GELU(tanh), residual"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "exmi cnn forward_kernel<64>",
                     "flop_per_launch": B * CNN_FLOP_PER_SAMPLE, "avg_launch_us": cnn_s * 1e6},
        "wall_s": {"corrected_coarse": wall_c, "plain_fine": wall_f, "plain_coarse": wall_p,
                   "training": t_train, "training_data": t_data},
        "learned_vs_plain_fine_speedup": wall_f / wall_c,
        "train_loss": hist,
    }
    # accuracy against the float64 fine oracle on the first trajectories
    from oracle import spectral as osp
    from oracle import stepper as ost

    nb = args.cpu_sample_batch or 16
    horizons = {k: k * DTC for k in (10, 20, 40) if k <= K}
    o64 = ost.KuramotoSivashinskyConservative(1, L, NF, DTF)
    u = u0[:nb].cpu().numpy().astype(np.float64)
    ref = {}
    t0 = time.perf_counter()
    for s_ in range(1, max(horizons) * ST + 1):
        u = o64.step_batch(u)
        if s_ % ST == 0 and s_ // ST in horizons:
            ref[s_ // ST] = np.stack([osp.map_between_resolutions(v, NC) for v in u])
    t_o = time.perf_counter() - t0
    trc, trf, trp = out["c"][:nb].cpu().numpy(), out["f"][:nb], out["p"][:nb].cpu().numpy()
    acc = {}
    ladder = {"plain_n128_dt05": (128, 0.5), "plain_n128_dt025": (128, 0.25)}
    lad_out, lad_wall = {}, {}
    for name, (n, dt) in ladder.items():
        stp = KS(1, L, n, dt)
        u0n = ex.map_between_resolutions(u0, n)
        steps = int(round(K * DTC / dt))
        stp.rollout_batch(u0n, 2)
        w_, _ = timed_rollout(lambda: out.__setitem__(name, stp.rollout_batch(u0n, steps)), sync)
        lad_wall[name] = w_
        lad_out[name] = (out[name][:nb], int(round(DTC / dt)))
        del out[name]
    for k, tphys in horizons.items():
        r = rel_l2_rows
        row = {"corrected_coarse": float(np.median(r(trc[:, k - 1], ref[k]))),
               "plain_coarse": float(np.median(r(trp[:, k - 1], ref[k]))),
               "plain_fine_gpu_f32": float(np.median(r(
                   ex.map_between_resolutions(trf[:, k * ST - 1].contiguous(), NC).cpu().numpy(), ref[k])))}
        for name, (tr_, per) in lad_out.items():
            row[name] = float(np.median(r(ex.map_between_resolutions(tr_[:, k * per - 1].contiguous(), NC)
                                          .cpu().numpy(), ref[k])))
        acc[f"t={tphys:g}"] = row
    res["median_rel_l2_vs_fp64_fine"] = acc
    res["ladder_wall_s"] = lad_wall
    res["nan_trajectories"] = {"corrected": int((~torch.isfinite(out["c"][:, -1]).all(-1).all(-1)).sum().item()),
                               "plain_coarse": int((~torch.isfinite(out["p"][:, -1]).all(-1).all(-1)).sum().item())}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = {"value": nb * max(horizons) * ST / t_o, "unit": "state-steps/s (fine, float64)",
                               "cores": os.cpu_count() or 1, "kind": "port",
                               "sample": f"NumPy/SciPy float64 fine stepper, {nb} trajectories x "
                                         f"{max(horizons) * ST} steps ({t_o:.1f} s)"}
    return res


def wl_cnn_forward(args, ex, _lib, torch):
    """The corrector network alone: exmi_cnn_forward on 8192 samples of 64
    points (seeded parameters and inputs), `--steps` launches timed with
    events on the launching stream. Its roofline is the BF16 matrix-core one."""
    from exponax_amd import corrector as cr

    B, K, W = args.batch or 8192, args.steps or 20, args.warmup
    cnn = cr.PeriodicCNN(seed=0)
    gen = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn(B, 1, 64, generator=gen).to("cuda:0")
    p = cnn.params.detach()
    for _ in range(max(W, 1)):
        cr.cnn_forward(p, x)
    torch.cuda.synchronize()
    out = {}

    def run():
        for _ in range(K):
            out["y"] = cr.cnn_forward(p, x)

    wall, gpu_s = timed_rollout(run, torch.cuda.synchronize)
    per = gpu_s / K
    achieved = B * CNN_FLOP_PER_SAMPLE / per / 1e12
    return {
        "metric": METRIC, "value": B * K / wall, "unit": "samples/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (f32 accumulate)", "data": "synthetic (seeded normal inputs, seeded parameters)",
        "config": {"workload": "cnn_forward_b8192_n64", "network": "5x Conv1d k5 32ch GELU(tanh), residual"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "exmi cnn forward_kernel<64>",
                     "flop_per_launch": B * CNN_FLOP_PER_SAMPLE, "avg_launch_us": per * 1e6},
    }


WORKLOADS = {"burgers1d": wl_burgers1d, "ks1d": wl_ks1d, "kolmogorov2d": wl_kolmogorov2d, "ks2d": wl_ks2d,
             "ks2d_grad": wl_ks2d_grad, "corrector": wl_corrector, "cnn_forward": wl_cnn_forward,
             "ns3d": wl_ns3d, "kolmogorov2d_513": wl_kolmogorov2d_513}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the workload's)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="burgers1d")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--strategy", type=int, default=0,
                    help="ks2d: force a rollout strategy (-1, 1, 2, 3; default: the autotuner's)")
    ap.add_argument("--cpu-sample-batch", type=int, default=0)
    ap.add_argument("--cpu-sample-steps", type=int, default=0)
    ap.add_argument("--train-iters", type=int, default=0, help="corrector workload: training iterations")
    ap.add_argument("--rollout-steps", type=int, default=0,
                    help="burgers1d: ETDRK steps per rollout graph (default 1000)")
    ap.add_argument("--no-physical-roundtrip", action="store_true",
                    help="burgers1d: skip the per-step physical round-trip rollout measured beside the default")
    ap.add_argument("--no-corrector-timing", action="store_true",
                    help="burgers1d: skip the learned-corrector inference timing added to the line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the trajectories of a batch "
              "shard across GPUs without communication, so one GPU's share is the measured unit)", file=sys.stderr)
        return 2

    import torch

    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 3}))
        return 0

    sys.path.insert(0, HERE)
    import exponax_amd as ex
    from exponax_amd import _lib

    res = WORKLOADS[args.workload](args, ex, _lib, torch)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
