"""Build the gfx950 shared library ``exponax_amd/_lib/libexmi.so`` in-tree.

Compiles every HIP translation unit under ``csrc/`` with hipcc for gfx950
(one object per 1-D transform size for the fused step kernels), in parallel,
and links them into one C-ABI shared library. Packed-FP32 instruction
selection is disabled for all device code.

Usage: ``python -m exponax_amd.build [-j JOBS]``.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
BUILDDIR = os.path.join(HERE, "_build")
LIBNAME = "libexmi.so"

# transform size M = N/2 -> points per lane (24 and 25: single-lane
# mixed-radix transforms for N = 48 and N = 50; odd M keeps the natural
# spectral layout, the paired one needs an even number of slots)
SIZES_1D = {8: 8, 16: 8, 24: 24, 25: 25, 32: 8, 64: 8, 128: 8, 256: 8, 512: 8, 1024: 16, 2048: 16}
# 2-D grid sizes N (N x N)
SIZES_2D = (32, 64, 128, 256, 512)


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _jobs():
    jobs = [("coef", os.path.join(CSRC, "coef.hip"), []), ("abi", os.path.join(CSRC, "abi.hip"), []),
            ("metrics", os.path.join(CSRC, "metrics.hip"), []),
            ("corrector", os.path.join(CSRC, "corrector.hip"), []),
            ("spectral3d", os.path.join(CSRC, "spectral3d.hip"), []),
            ("gfft", os.path.join(CSRC, "gfft.hip"), []),
            ("ic", os.path.join(CSRC, "ic.hip"), [])]
    for m, pt in SIZES_1D.items():
        defs = [f"-DEXMI_M={m}", f"-DEXMI_PT={pt}"] + (["-DEXMI_LAYOUT_1D=NATURAL"] if m % 2 else [])
        jobs.append((f"step1d_{m}", os.path.join(CSRC, "step1d_inst.hip"), defs))
    for n in SIZES_2D:
        jobs.append((f"step2d_{n}", os.path.join(CSRC, "step2d_inst.hip"), [f"-DEXMI_N2={n}"]))
    # longest compiles first
    jobs.sort(key=lambda j: 0 if j[0] in ("step1d_1024", "step1d_2048", "step1d_512") else 1)
    return jobs


def _newest_source() -> float:
    return max(os.path.getmtime(os.path.join(CSRC, f)) for f in os.listdir(CSRC)) if os.path.isdir(CSRC) else 0.0


def _compile(job, force=True):
    name, src, defs = job
    obj = os.path.join(BUILDDIR, name + ".o")
    if not force and os.path.exists(obj) and os.path.getmtime(obj) > max(
            _newest_source(), os.path.getmtime(os.path.join(os.path.dirname(HERE), "include", "exmi.h"))):
        return obj, 0.0
    cmd = [hipcc()] + COMMON + defs + ["-c", src, "-o", obj]
    t = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile {name} failed:\n{' '.join(cmd)}\n{p.stderr[-6000:]}")
    return obj, time.time() - t


def build(jobs: int | None = None, verbose: bool = False, force: bool = True) -> str:
    """Compile and link. With force=False, objects newer than every source
    are reused (development convenience); the default recompiles all."""
    os.makedirs(BUILDDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, dt in ex.map(lambda j: _compile(j, force), _jobs()):
            objs.append(obj)
            if verbose:
                print(f"  {os.path.basename(obj)}: {dt:.1f}s", flush=True)
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr[-4000:]}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--incremental", action="store_true", help="reuse objects newer than all sources")
    args = ap.parse_args()
    t0 = time.time()
    path = build(args.j, verbose=True, force=not args.incremental)
    print(f"built {path} in {time.time() - t0:.1f}s")
    sys.exit(0)
