"""ctypes binding of the gfx950 C-ABI library ``_lib/libexmi.so``.

The library is the only compute path: when it is missing, or no HIP device
is visible, every entry point raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libexmi.so")

EXMI_NL_ZERO = 0
EXMI_NL_CONVECTION = 1
EXMI_NL_GRADIENT_NORM = 2
EXMI_NL_POLYNOMIAL = 3
EXMI_NL_GENERAL = 4
EXMI_NL_VORTICITY = 5
EXMI_NL_KOLMOGOROV = 6
EXMI_NL_CAHN_HILLIARD = 7
EXMI_NL_GRAY_SCOTT = 8
EXMI_NL_BELOUSOV_ZHABOTINSKY = 9
EXMI_NL_PROJECTED_CONVECTION = 10
EXMI_NL_PROJECTED_CONVECTION_KOLMOGOROV = 11
EXMI_MAX_POLY = 8
EXMI_FLAG_EXACT_FORM = 1
EXMI_FLAG_FOURIER_RESIDENT = 2
EXMI_FLAG_PHYSICAL_ROLLOUT = 4

EXMI_ERR_INVALID = -1
EXMI_ERR_HIP = -2
EXMI_ERR_ALLOC = -3
EXMI_ERR_UNSUPPORTED = -4

# every symbol include/exmi.h declares
EXPORTED = (
    "exmi_plan_create", "exmi_plan_destroy", "exmi_step", "exmi_step_fourier", "exmi_nonlinear",
    "exmi_rollout", "exmi_rollout_prepare", "exmi_rollout_group", "exmi_repeat", "exmi_rfft", "exmi_irfft", "exmi_plan_num_coefs", "exmi_plan_get_tables",
    "exmi_plan_rollout_strategy", "exmi_plan_set_rollout_strategy", "exmi_last_error", "exmi_info",
    "exmi_step_vjp", "exmi_rollout_vjp", "exmi_nrmse", "exmi_spectrum", "exmi_spectrum_binned",
    "exmi_map_resolution", "exmi_ic_tfs", "exmi_wave_step_fourier", "exmi_poisson_fourier", "exmi_axpy",
    "exmi_cnn_workspace_bytes", "exmi_cnn_forward", "exmi_cnn_backward", "exmi_corrected_rollout",
    "exmi_irfft_noalloc", "exmi_rollout_checked", "exmi_first_bad_step", "exmi_plan_fft_config",
    "exmi_plan_set_fft_config",
)
EXMI_NOALLOC_SCRATCH_BYTES = 32 << 20
EXMI_CNN_NUM_PARAMS = 15809


class PlanDesc(C.Structure):
    """Mirror of ``exmi_plan_desc`` in include/exmi.h."""

    _fields_ = [
        ("num_spatial_dims", C.c_int32),
        ("num_points", C.c_int32),
        ("num_channels", C.c_int32),
        ("batch", C.c_int32),
        ("order", C.c_int32),
        ("nonlin_kind", C.c_int32),
        ("conv_single_channel", C.c_int32),
        ("conv_conservative", C.c_int32),
        ("zero_mode_fix", C.c_int32),
        ("num_poly", C.c_int32),
        ("injection_mode", C.c_int32),
        ("lin_batch", C.c_int32),
        ("lin_channels", C.c_int32),
        ("num_circle_points", C.c_int32),
        ("flags", C.c_int32),
        ("domain_extent", C.c_double),
        ("dt", C.c_double),
        ("circle_radius", C.c_double),
        ("dealiasing_fraction", C.c_double),
        ("scale", C.c_double),
        ("general_scales", C.c_double * 3),
        ("poly", C.c_double * EXMI_MAX_POLY),
        ("injection_scale", C.c_double),
        ("host_linear_operator", C.POINTER(C.c_double)),
        ("host_linear_operator_imag", C.POINTER(C.c_double)),
        ("reaction_rates", C.c_double * 2),
        ("host_nonlin_factor", C.POINTER(C.c_double)),
    ]


class ExmiError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()


def _declare(lib):
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    lib.exmi_plan_create.argtypes = [C.POINTER(PlanDesc), C.POINTER(vp)]
    lib.exmi_plan_destroy.argtypes = [vp]
    for name in ("exmi_step", "exmi_step_fourier", "exmi_nonlinear"):
        getattr(lib, name).argtypes = [vp, vp, i64, vp, i64, vp]
    lib.exmi_rollout.argtypes = [vp, vp, i64, vp, i32, i32, i32, vp]
    lib.exmi_rollout_prepare.argtypes = [vp, vp, i64, vp, i32, i32]
    lib.exmi_rollout_group.argtypes = [C.POINTER(vp), i32, C.POINTER(vp), C.POINTER(i64), C.POINTER(vp), i32, i32,
                                       i32, vp]
    lib.exmi_repeat.argtypes = [vp, vp, i64, vp, i32, i32, vp]
    lib.exmi_rfft.argtypes = [i32, i32, i32, vp, vp, vp]
    lib.exmi_irfft.argtypes = [i32, i32, i32, vp, vp, vp]
    lib.exmi_irfft_noalloc.argtypes = [i32, i32, i32, vp, vp, vp]
    lib.exmi_rollout_checked.argtypes = [vp, vp, i64, vp, i32, i32, C.c_float, vp, vp]
    lib.exmi_first_bad_step.argtypes = [i32, i32, i64, vp, i64, i64, C.c_float, vp, vp]
    lib.exmi_plan_num_coefs.argtypes = [vp]
    lib.exmi_plan_get_tables.argtypes = [vp, vp, i64]
    lib.exmi_plan_rollout_strategy.argtypes = [vp]
    lib.exmi_plan_set_rollout_strategy.argtypes = [vp, i32]
    lib.exmi_plan_fft_config.argtypes = [vp, vp, i32]
    lib.exmi_plan_set_fft_config.argtypes = [vp, i32]
    lib.exmi_step_vjp.argtypes = [vp, vp, i64, vp, i64, vp, i64, vp]
    lib.exmi_rollout_vjp.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp, i32, vp]
    lib.exmi_nrmse.argtypes = [i32, i32, i64, vp, i64, vp, i64, vp, vp]
    lib.exmi_spectrum.argtypes = [i32, i32, i32, vp, vp, i32, vp]
    lib.exmi_spectrum_binned.argtypes = [i32, i32, i32, vp, vp, i32, i32, vp]
    lib.exmi_ic_tfs.argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp]
    lib.exmi_wave_step_fourier.argtypes = [i32, i32, i32, C.c_double, C.c_double, C.c_double, vp, vp]
    lib.exmi_poisson_fourier.argtypes = [i32, i32, i32, C.c_double, i32, vp, vp]
    lib.exmi_axpy.argtypes = [C.c_int64, C.c_float, vp, vp, vp, vp]
    lib.exmi_map_resolution.argtypes = [i32, i32, i32, i32, vp, vp, i32, vp]
    lib.exmi_cnn_workspace_bytes.argtypes = [i32, i32]
    lib.exmi_cnn_forward.argtypes = [vp, i32, i32, vp, vp, vp, vp]
    lib.exmi_cnn_backward.argtypes = [vp, i32, i32, vp, vp, vp, vp, vp]
    lib.exmi_corrected_rollout.argtypes = [vp, vp, vp, i64, vp, i32, i32, vp]
    lib.exmi_last_error.restype = C.c_char_p
    lib.exmi_info.argtypes = [C.c_char_p, i32]
    for name in EXPORTED:
        if name != "exmi_last_error":
            getattr(lib, name).restype = C.c_int
    lib.exmi_cnn_workspace_bytes.restype = i64


def load_library(path: str = LIB_PATH):
    """Load (once) and return the ctypes handle. Raises if the .so is absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise ExmiError(
                    f"exponax_amd native library not built: {path} is missing "
                    "(run `python -m exponax_amd.build`)")
            lib = C.CDLL(path)
            _declare(lib)
            _lib = lib
    return _lib


def lib():
    """The library, after checking that a HIP device is present."""
    L = load_library()
    _require_device()
    return L


_device_checked = False


def _require_device():
    global _device_checked
    if _device_checked:
        return
    import torch

    if not torch.cuda.is_available():
        raise ExmiError("exponax_amd needs a HIP device (torch.cuda.is_available() is False); "
                        "there is no CPU fallback")
    _device_checked = True


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = _lib.exmi_last_error().decode() if _lib is not None else ""
        if rc == EXMI_ERR_INVALID:
            raise ValueError(msg or what)
        if rc == EXMI_ERR_UNSUPPORTED:
            raise NotImplementedError(msg or what)
        raise ExmiError(f"{what}: {msg} (status {rc})")


def stream_ptr(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


class Plan:
    """Owns one ``exmi_plan``; created from a PlanDesc plus the L table
    (float64, or complex128 for odd-derivative operators)."""

    def __init__(self, desc: PlanDesc, linear_operator: np.ndarray, nonlin_factor=None):
        L = lib()
        self._F = None
        if nonlin_factor is not None:
            self._F = np.ascontiguousarray(nonlin_factor, dtype=np.float64)
            desc.host_nonlin_factor = self._F.ctypes.data_as(C.POINTER(C.c_double))
        lin = np.asarray(linear_operator)
        self._L = np.ascontiguousarray(lin.real, dtype=np.float64)
        desc.host_linear_operator = self._L.ctypes.data_as(C.POINTER(C.c_double))
        self._Li = None
        if np.iscomplexobj(lin):
            self._Li = np.ascontiguousarray(lin.imag, dtype=np.float64)
            desc.host_linear_operator_imag = self._Li.ctypes.data_as(C.POINTER(C.c_double))
        self.desc = desc
        h = C.c_void_p()
        check(L.exmi_plan_create(C.byref(desc), C.byref(h)), "exmi_plan_create")
        self._h = h
        self._lib = L

    @property
    def handle(self):
        return self._h

    def num_coefs(self) -> int:
        return int(self._lib.exmi_plan_num_coefs(self._h))

    def tables(self) -> np.ndarray:
        d = self.desc
        nr = (d.num_points ** (d.num_spatial_dims - 1)) * (d.num_points // 2 + 1)
        n = d.lin_batch * d.lin_channels * self.num_coefs() * nr
        out = np.empty(n, dtype=np.float32)
        check(self._lib.exmi_plan_get_tables(self._h, out.ctypes.data, n), "exmi_plan_get_tables")
        return out.reshape(d.lin_batch, d.lin_channels, self.num_coefs(), nr)

    def rollout_strategy(self) -> int:
        return int(self._lib.exmi_plan_rollout_strategy(self._h))

    def set_rollout_strategy(self, strategy: int) -> None:
        check(self._lib.exmi_plan_set_rollout_strategy(self._h, int(strategy)), "exmi_plan_set_rollout_strategy")

    def fft_config(self) -> dict | None:
        """The runtime-size transforms' knobs (None for plans on the fused
        kernels): code = reg * 100000000 + max_radix * 100000 + points per
        workgroup (reg 1: radices 9..17 and 19 as register butterflies), where it
        came from (default / timed / cache / set), lines per workgroup and
        the pass radices."""
        buf = (C.c_int32 * 32)()
        n = self._lib.exmi_plan_fft_config(self._h, buf, 32)
        check(min(n, 0), "exmi_plan_fft_config")
        if n == 0:
            return None
        v = list(buf[:n])
        return {"code": v[0], "source": ("default", "timed", "cache", "set")[v[1]], "lines_per_workgroup": v[2],
                "radices": v[4:4 + v[3]]}

    def set_fft_config(self, code: int) -> None:
        check(self._lib.exmi_plan_set_fft_config(self._h, int(code)), "exmi_plan_set_fft_config")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.exmi_plan_destroy(h)
            except Exception:
                pass
            self._h = None
