"""exponax_amd — exponax's ETDRK pseudo-spectral step on AMD MI355X (gfx950).

The stepper constructors, ``(C, N, ..., N)`` state layout, ``__call__`` /
``step_fourier`` / ``rollout`` / ``repeat`` follow exponax 0.2.0; every step
runs in hand-written HIP kernels behind the C ABI in ``include/exmi.h``.
"""
from . import autograd, corrector, ic, io, metrics, nonlin_fun, spectral, stepper
from ._base_stepper import BaseStepper, Ensemble, ensemble
from ._lib import ExmiError, load_library
from ._forced_stepper import ForcedStepper
from ._poisson import Poisson
from ._repeated_stepper import RepeatedStepper
from ._utils import make_grid, repeat, rollout, rollout_group, vmap
from .spectral import fft, get_spectrum, ifft, map_between_resolutions

__all__ = ["BaseStepper", "Ensemble", "ForcedStepper", "Poisson", "RepeatedStepper", "ensemble", "ExmiError", "autograd", "corrector", "fft", "get_spectrum", "ifft", "io",
           "load_library", "make_grid", "map_between_resolutions", "metrics", "nonlin_fun", "repeat", "rollout",
           "rollout_group", "spectral", "stepper", "vmap"]
