"""CPU oracle for the ETDRK pseudo-spectral step — TEST INFRASTRUCTURE ONLY.

A NumPy/SciPy restatement of exponax 0.2.0's hot path (``_spectral.py``,
``nonlin_fun/``, ``etdrk/``, ``_base_stepper.py``, the config steppers,
``_utils.py`` rollout/repeat, ``metrics/_spatial.py`` nRMSE), in float64 or
float32. Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import it, and only as the checker / CPU baseline; the
product package ``exponax_amd`` never imports it.

Pinning (each function's check lives in tests/test_oracle_*.py):
- spectral grids/operators/masks: closed forms and the reference's literal
  mask arrays (tests/golden/filter_masks.json);
- fft/ifft: brute-force O(N^2) DFT and numpy.fft;
- nonlinear functions: closed forms on sines (reference
  tests/test_nonlinear_funs.py), conservative == non-conservative, DC
  fixes, Kolmogorov injection value;
- ETDRK coefficients: closed-form phi functions in mpmath; ETDRK0 exactness;
  convergence orders 1-4 (reference tests/test_etdrk.py);
- steppers: equivalences and invariants of reference
  tests/test_builtin_solvers.py; rollout == loop.
Bit-level agreement with JAX's XLA FFT is not pinnable here (JAX is absent).
"""
from . import adjoint, etdrk, ic, nonlin, spectral, stepper  # noqa: F401
