"""Throughput benchmark of the fused ETDRK step on one MI355X.

Workload (the batched Burgers rollout): Burgers 1-D, N=1024, L=1, dt=1e-3,
ETDRK2, 4096 trajectories advanced together, per-trajectory viscosity
nu log-uniform in [1e-4, 1e-2], float32 truncated-Fourier ICs (cutoff 8,
seed 2), rollout of --steps steps captured in one hipGraph, every state
written to the trajectory buffer (rollout semantics).

A bench step is one ETDRK2 step of all 4096 trajectories; `value` is
state-steps/s = 4096 * steps / wall time of the timed graph replay, with the
ICs already resident in HBM. The rollout runs as the strategy the library's
autotuner picks (one multi-step launch, or per-step launches). The roofline
uses the algorithmic bytes of one launch: 4096 * steps-per-launch *
(4*1024 read + 4*1024 write + 12*513 operator tables) B.

The CPU baseline (rank 0 only) runs the NumPy/SciPy FP32 restatement of the
same step on a bounded sample (all 4096 trajectories for 300 steps) and
doubles as the parity check of the GPU trajectory at steps 1 and 10.
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
N, L, DT = 1024, 1.0, 1e-3
CUTOFF, SEED = 8, 2
HERE = os.path.dirname(os.path.abspath(__file__))


def truncated_fourier_ics(batch, n, length, cutoff, rng):
    """The TFS recipe: modes k = 1..cutoff, a ~ N(0,1), phase ~ U[0, 2pi),
    u = sum a cos(2 pi k x / L + phase), u /= max|u| (per trajectory, drawn
    in trajectory order from one stream; identical to the test fixtures)."""
    x = np.linspace(0.0, length, n, endpoint=False)
    k = np.arange(1, cutoff + 1, dtype=np.float64)
    out = np.empty((batch, 1, n))
    for b in range(batch):
        a = rng.standard_normal(cutoff)
        ph = rng.uniform(0.0, 2 * np.pi, cutoff)
        u = (a[:, None] * np.cos(2 * np.pi * k[:, None] * x[None, :] / length + ph[:, None])).sum(0)
        out[b, 0] = u / np.max(np.abs(u))
    return out


def read_traffic(workload, steps_per_launch):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (measured bytes per state-step, scaled to this run's launch)."""
    path = os.path.join(HERE, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        per_step = json.load(f).get("bytes_per_state_step")
    return None if per_step is None else per_step * 4096 * steps_per_launch


def cpu_baseline(u0, nus, gpu_traj, sample_b, sample_steps):
    """Time the FP32 oracle on a bounded sample and check the GPU trajectory."""
    sys.path.insert(0, HERE)
    from oracle import spectral as osp
    from oracle import stepper as ost

    cores = os.cpu_count() or 1
    osp.set_fft_workers(cores)
    members = [ost.Burgers(1, L, N, DT, diffusivity=float(nu), dtype=np.float32) for nu in nus[:sample_b]]
    ens = ost.Ensemble(members, dtype=np.float32, order=2)
    u = u0[:sample_b].astype(np.float32)
    checks = {}
    t0 = time.perf_counter()
    with np.errstate(all="ignore"):  # the lowest-viscosity trajectories blow up, as in the reference
        for s in range(1, sample_steps + 1):
            u = ens.step_batch(u)
            if s in (1, 10):
                checks[s] = u.copy()
    dt = time.perf_counter() - t0
    parity = {}
    for s, ref in checks.items():
        g = gpu_traj[s]
        ok = np.isfinite(ref).all(axis=(1, 2)) & np.isfinite(g).all(axis=(1, 2))
        num = np.sqrt(((g[ok].astype(np.float64) - ref[ok]) ** 2).sum(axis=(1, 2)))
        den = np.sqrt((ref[ok].astype(np.float64) ** 2).sum(axis=(1, 2)))
        r = num / den
        parity[f"rel_l2_step{s}_max"] = float(r.max())
        parity[f"rel_l2_step{s}_mean"] = float(r.mean())
    parity["checked_trajectories"] = int(sample_b)
    base = {"value": sample_b * sample_steps / dt, "unit": "state-steps/s", "cores": cores, "kind": "port",
            "sample": f"NumPy/SciPy FP32 restatement, first {sample_b} trajectories x {sample_steps} steps "
                      f"({dt:.1f} s), scipy.fft workers={cores}"}
    return base, parity


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-batch", type=int, default=4096)
    ap.add_argument("--cpu-sample-steps", type=int, default=300)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (trajectories shard across GPUs "
              "without communication, see DESIGN.md)", file=sys.stderr)
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

    B, K, W = args.batch, args.steps, args.warmup
    rng = np.random.default_rng(SEED)
    u0 = truncated_fourier_ics(B, N, L, CUTOFF, rng)
    nus = 10.0 ** rng.uniform(-4.0, -2.0, B)
    ens = ex.ensemble(lambda nu: ex.stepper.Burgers(1, L, N, DT, diffusivity=float(nu)), nus)
    plan = ens.plan(B)
    lib = _lib.lib()
    dev = torch.device("cuda:0")
    u0_d = torch.tensor(u0, dtype=torch.float32, device=dev)
    sz = N
    stream = _lib.stream_ptr()

    # warmup: W steps through a captured graph of their own
    if W > 0:
        tw = torch.empty((B, W, 1, N), dtype=torch.float32, device=dev)
        _lib.check(lib.exmi_rollout(plan.handle, u0_d.data_ptr(), sz, tw.data_ptr(), W, 0, 1, stream))
        torch.cuda.synchronize()
        del tw
    traj = torch.empty((B, K, 1, N), dtype=torch.float32, device=dev)
    _lib.check(lib.exmi_rollout_prepare(plan.handle, u0_d.data_ptr(), sz, traj.data_ptr(), K, 0))
    torch.cuda.synchronize()

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    _lib.check(lib.exmi_rollout(plan.handle, u0_d.data_ptr(), sz, traj.data_ptr(), K, 0, 1, stream))
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)

    value = B * K / wall
    strategy = plan.rollout_strategy()
    # per-step launches (strategy -1) or one multi-step launch of K steps
    n_launches = K if strategy < 0 else 1
    steps_per_launch = K // n_launches
    per_launch_s = gpu_ms / 1e3 / n_launches
    bytes_per_state_step = 4 * N + 4 * N + 12 * (N // 2 + 1)
    alg_bytes = B * steps_per_launch * bytes_per_state_step
    achieved = alg_bytes / per_launch_s / 1e9
    kernel = ("step1d_kernel<512,8,NL_CONV_CONS,2>" if strategy < 0 else
              f"roll1d_kernel<512,8,NL_CONV_CONS,2,{strategy}> ({steps_per_launch} steps per launch)")
    final = traj[:, K - 1].float()
    nan_traj = int((~torch.isfinite(final).all(dim=-1).all(dim=-1)).sum().item())

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "state-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded truncated-Fourier ICs, log-uniform viscosities)",
        "config": {"workload": "burgers1d_n1024_b4096_etdrk2_graph", "equation": "Burgers", "num_points": N,
                   "domain_extent": L, "dt": DT, "order": 2, "batch": B, "rollout_steps": K,
                   "per_trajectory_viscosity": "log-uniform [1e-4, 1e-2]", "ic": "TFS cutoff 8 seed 2",
                   "hipgraph": True},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic("burgers1d_n1024_b4096", steps_per_launch),
                     "alg_bytes_per_launch": alg_bytes, "alg_bytes_per_state_step": bytes_per_state_step,
                     "avg_launch_us": per_launch_s * 1e6, "launches": n_launches, "kernel": kernel},
        "rollout_strategy": strategy,
        "nan_trajectories": nan_traj,
    }
    if not args.no_cpu_baseline:
        nb = min(args.cpu_sample_batch, B)
        gsel = {s: traj[:nb, s - 1].cpu().numpy() for s in (1, 10) if s <= K}
        base, parity = cpu_baseline(u0, nus, gsel, nb, args.cpu_sample_steps)
        result["cpu_baseline"] = base
        result["parity"] = parity
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
