"""The plan/ABI layer (csrc/abi.hip: descriptor validation, plan set-up and
tear-down, error reporting) built for the host with AddressSanitizer and
UndefinedBehaviorSanitizer and driven through its validation and error
paths on the CPU. The device code is compiled unchanged (the sanitizer flags
apply to the host pass only); without a HIP device every plan set-up fails
at its first device call, which exercises the clean-up paths as well. The
rest of the library links from the regular build's objects."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "exponax_amd", "_build")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang"

DRIVER = r"""
#include <stdio.h>
#include <string.h>
#include "exmi.h"

static int n_checks = 0;
static void report(const char* what, int rc) {
  printf("%s: rc=%d msg=%s\n", what, rc, exmi_last_error());
  ++n_checks;
}

static exmi_plan_desc base(int D, int N) {
  exmi_plan_desc d;
  memset(&d, 0, sizeof d);
  d.num_spatial_dims = D;
  d.num_points = N;
  d.num_channels = 1;
  d.batch = 2;
  d.order = 2;
  d.nonlin_kind = EXMI_NL_GRADIENT_NORM;
  d.lin_batch = 1;
  d.lin_channels = 1;
  d.num_circle_points = 16;
  d.domain_extent = 10.0;
  d.dt = 0.1;
  d.circle_radius = 1.0;
  d.dealiasing_fraction = 2.0 / 3.0;
  d.scale = 1.0;
  return d;
}

int main(void) {
  static double L[3 * 256 * 256];
  exmi_plan* p = 0;
  exmi_plan_desc d;
  report("null desc", exmi_plan_create(0, &p));
  d = base(4, 64); d.host_linear_operator = L; report("D=4", exmi_plan_create(&d, &p));
  d = base(1, 5000); d.host_linear_operator = L; report("1-D N=5000", exmi_plan_create(&d, &p));
  d = base(2, 1); d.host_linear_operator = L; report("2-D N=1", exmi_plan_create(&d, &p));
  d = base(3, 4097); d.host_linear_operator = L; report("3-D N=4097", exmi_plan_create(&d, &p));
  d = base(1, 64); d.host_linear_operator = L; d.order = 5; report("order 5", exmi_plan_create(&d, &p));
  d = base(1, 64); d.host_linear_operator = L; d.lin_batch = 3; report("lin_batch", exmi_plan_create(&d, &p));
  d = base(1, 64); report("no L", exmi_plan_create(&d, &p));
  d = base(1, 64); d.host_linear_operator = L;
  d.flags = EXMI_FLAG_FOURIER_RESIDENT | EXMI_FLAG_PHYSICAL_ROLLOUT; report("flags", exmi_plan_create(&d, &p));
  d = base(2, 64); d.host_linear_operator = L; d.nonlin_kind = EXMI_NL_CONVECTION; d.num_channels = 3;
  report("2-D convection C=3", exmi_plan_create(&d, &p));
  d = base(1, 64); d.host_linear_operator = L; d.num_poly = 9; d.nonlin_kind = EXMI_NL_POLYNOMIAL;
  report("polynomial 9", exmi_plan_create(&d, &p));
  /* valid descriptors: set-up stops at the first device call */
  d = base(1, 256); d.host_linear_operator = L; report("valid 1-D", exmi_plan_create(&d, &p));
  d = base(2, 64); d.host_linear_operator = L; report("valid 2-D", exmi_plan_create(&d, &p));
  d = base(3, 16); d.host_linear_operator = L; report("valid 3-D", exmi_plan_create(&d, &p));
  report("destroy null", exmi_plan_destroy(0));
  report("step null", exmi_step(0, 0, 0, 0, 0, 0));
  report("step_fourier null", exmi_step_fourier(0, 0, 0, 0, 0, 0));
  report("rollout null", exmi_rollout(0, 0, 0, 0, 1, 0, 1, 0));
  report("rollout_checked null", exmi_rollout_checked(0, 0, 0, 0, 1, 0, 0.f, 0, 0));
  report("repeat null", exmi_repeat(0, 0, 0, 0, 1, 1, 0));
  report("strategy null", exmi_plan_set_rollout_strategy(0, 1));
  report("rfft D=4", exmi_rfft(4, 64, 1, 0, 0, 0));
  report("rfft N=5000", exmi_rfft(1, 5000, 1, 0, 0, 0));
  report("irfft 2-D N=1", exmi_irfft(2, 1, 1, 0, 0, 0));
  d = base(2, 100); d.host_linear_operator = L; report("valid 2-D N=100", exmi_plan_create(&d, &p));
  d = base(1, 11); d.host_linear_operator = L; report("valid 1-D N=11", exmi_plan_create(&d, &p));
  report("fft_config null", exmi_plan_fft_config(0, 0, 0));
  report("set_fft_config null", exmi_plan_set_fft_config(0, 800512));
  report("irfft_noalloc", exmi_irfft_noalloc(2, 64, 1, 0, 0, 0));
  report("nrmse null", exmi_nrmse(1, 1, 8, 0, 8, 0, 8, 0, 0));
  report("first_bad null", exmi_first_bad_step(1, 1, 8, 0, 8, 8, 0.f, 0, 0));
  report("poisson odd order", exmi_poisson_fourier(1, 64, 1, 1.0, 3, (float*)L, 0));
  report("map_resolution D=5", exmi_map_resolution(5, 64, 32, 1, 0, 0, 0, 0));
  char info[256];
  report("info", exmi_info(info, (int)sizeof info));
  printf("checks=%d\n", n_checks);
  return 0;
}
"""


def test_abi_layer_under_address_and_ub_sanitizers(tmp_path):
    objs = [o for o in glob.glob(os.path.join(BUILD, "*.o")) if os.path.basename(o) != "abi.o"]
    if not objs or not os.path.exists(HIPCC) or not os.path.exists(CLANG):
        pytest.skip("needs the library's build objects (exponax_amd/_build) and the ROCm compilers")
    obj = str(tmp_path / "abi_san.o")
    cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Wno-unused-command-line-argument",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all",
           "-c", os.path.join(ROOT, "exponax_amd", "csrc", "abi.hip"), "-o", obj]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    lib = str(tmp_path / "libexmi_san.so")
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", lib, obj] + objs, check=True,
                   capture_output=True, text=True)
    src = tmp_path / "driver.c"
    src.write_text(DRIVER)
    exe = str(tmp_path / "driver")
    subprocess.run([CLANG, "-fsanitize=address,undefined", "-g", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", str(tmp_path), "-lexmi_san", f"-Wl,-rpath,{tmp_path}", "-o", exe], check=True,
                   capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    p = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    rc = {k: int(v.split()[0]) for k, v in (ln.split(": rc=", 1) for ln in p.stdout.splitlines() if ": rc=" in ln)}
    msg = {k: v.split("msg=", 1)[1] for k, v in (ln.split(": rc=", 1) for ln in p.stdout.splitlines() if ": rc=" in ln)}
    invalid, hip, unsupported = -1, -2, -4
    expected = {
        "null desc": invalid, "D=4": unsupported, "1-D N=5000": unsupported, "2-D N=1": unsupported,
        "3-D N=4097": unsupported, "order 5": unsupported, "lin_batch": invalid, "no L": invalid, "flags": invalid,
        "2-D convection C=3": invalid, "polynomial 9": unsupported,
        # valid descriptors: no device here, so the set-up fails at its first HIP call and cleans up
        "valid 1-D": hip, "valid 2-D": hip, "valid 3-D": hip, "irfft_noalloc": hip,
        "destroy null": 0, "step null": invalid, "step_fourier null": invalid, "rollout null": invalid,
        "rollout_checked null": invalid, "repeat null": invalid, "strategy null": invalid,
        "rfft D=4": unsupported, "rfft N=5000": unsupported, "irfft 2-D N=1": unsupported,
        "valid 2-D N=100": hip, "valid 1-D N=11": hip, "fft_config null": invalid, "set_fft_config null": invalid,
        "nrmse null": invalid, "first_bad null": invalid, "poisson odd order": invalid,
        "map_resolution D=5": unsupported, "info": 0,
    }
    assert rc == expected, rc
    assert "supported sizes are 2..4096" in msg["1-D N=5000"] and "lin_batch" in msg["lin_batch"]
    assert "exclude each other" in msg["flags"]
    assert "checks=34" in p.stdout
    shutil.rmtree(tmp_path, ignore_errors=True)
