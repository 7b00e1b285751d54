/*
 * exmi.h — C ABI of the MI355X (gfx950) ETDRK pseudo-spectral stepper.
 *
 * Drop-in boundary for exponax 0.2.0's hot path: every entry point below
 * replaces one reference interface (cited as file:line of exponax/). All
 * pointers are device pointers unless named host_*; all calls are
 * asynchronous on the given HIP stream (passed as void*; NULL = the default
 * stream). Every function returns 0 (EXMI_OK) or a negative status, and
 * exmi_last_error() returns a thread-local message for the last failure.
 * No C++ exception crosses this boundary.
 *
 * State layout (reference _base_stepper.py:241-271, vmapped over a leading
 * batch axis): float32 (B, C, N[, N]) row-major; spectra are complex64
 * (B, C, [N,] N/2+1) with interleaved (re, im) floats, exactly
 * jnp.fft.rfftn's layout (_spectral.py:614-656).
 *
 * Threading: a plan is not re-entrant across streams; calls that share a
 * plan must be ordered on one stream.
 */
#ifndef EXMI_H
#define EXMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum exmi_status {
  EXMI_OK = 0,
  EXMI_ERR_INVALID = -1,     /* bad argument / unsupported configuration   */
  EXMI_ERR_HIP = -2,         /* a HIP runtime call failed                  */
  EXMI_ERR_ALLOC = -3,       /* device allocation failed                   */
  EXMI_ERR_UNSUPPORTED = -4  /* valid for the reference, not built here    */
};

/* Nonlinear term kinds (reference nonlin_fun/__init__.py:34-62). */
enum exmi_nonlin_kind {
  EXMI_NL_ZERO = 0,        /* ZeroNonlinearFun        nonlin_fun/_zero.py:7            */
  EXMI_NL_CONVECTION = 1,  /* ConvectionNonlinearFun  nonlin_fun/_convection.py:7      */
  EXMI_NL_GRADIENT_NORM = 2, /* GradientNormNonlinearFun nonlin_fun/_gradient_norm.py:8 */
  EXMI_NL_POLYNOMIAL = 3,  /* PolynomialNonlinearFun  nonlin_fun/_polynomial.py:6      */
  EXMI_NL_GENERAL = 4,     /* GeneralNonlinearFun     nonlin_fun/_general_nonlinear.py:9 */
  EXMI_NL_VORTICITY = 5,   /* VorticityConvection2d   nonlin_fun/_vorticity_convection.py:8 */
  EXMI_NL_KOLMOGOROV = 6,  /* VorticityConvection2dKolmogorov  ..._vorticity_convection.py:102 */
  /* scale * Laplace F[u^3]: CahnHilliardNonlinearFun
   * stepper/reaction/_cahn_hilliard.py:10-41 (scale = nu c3) */
  EXMI_NL_CAHN_HILLIARD = 7,
  /* Channel-coupled reactions (two and three channels), one per-channel
   * linear operator row each (lin_channels = num_channels):
   * GrayScottNonlinearFun stepper/reaction/_gray_scott.py:9-49 (feed and kill
   * rates in reaction_rates), BelousovZhabotinskyNonlinearFun
   * stepper/reaction/_belousov_zhabotinsky.py:9-47. 1-D plans run every
   * step, rollout and repeat through the multi-step kernel with the channels
   * of a trajectory in one wave (N <= 1024; no exmi_step_fourier /
   * exmi_nonlinear); 2-D plans couple them in the row kernel. */
  EXMI_NL_GRAY_SCOTT = 8,
  EXMI_NL_BELOUSOV_ZHABOTINSKY = 9,
  /* P(u x omega), three channels, 3-D plans: ProjectedConvection3d
   * nonlin_fun/_projected_convection.py:19-136 (Leray projection
   * nonlin_fun/_leray.py:86-103); the Kolmogorov variant (:139-226) adds
   * injection_scale * coef_extraction scaling at (0, injection_mode, 0) of
   * the first channel */
  EXMI_NL_PROJECTED_CONVECTION = 10,
  EXMI_NL_PROJECTED_CONVECTION_KOLMOGOROV = 11
};

#define EXMI_MAX_POLY 8

/* Plan flags.
 * EXMI_FLAG_EXACT_FORM: evaluate the non-conservative convection u u_x with
 * its two inverse transforms as the reference does
 * (nonlin_fun/_convection.py:140-171), even when the dealiasing is alias-free
 * for quadratic products (3 * kmax < N). Without the flag such plans evaluate
 * (1/2) d/dx (u^2) instead: on every retained mode both are the same exact
 * (alias-free) convolution, so only the float32 rounding differs, and one
 * inverse transform per nonlinear evaluation is saved. */
#define EXMI_FLAG_EXACT_FORM 1

/* Rollout forms of 1-D ETDRK2 plans. The reference's rollout
 * (_utils.py:167-172 over _base_stepper.py:201-220) transforms every state
 * forward, steps, and transforms back. When every spectrum the step produces
 * has real DC and Nyquist bins (a real L, and a nonlinear term whose Nyquist
 * mode the dealias mask removes), the forward transform of an emitted state
 * returns the spectrum it came from exactly in exact arithmetic; for such
 * plans rollouts and repeats keep each trajectory's spectrum on chip between
 * steps and write its inverse transform ("Fourier-resident"). Results differ
 * from the per-step round trip by float32 rounding only, and one forward
 * transform per step is saved. This is the default for those plans; the
 * per-step launches (exmi_plan_set_rollout_strategy(-1)) keep the round trip.
 * EXMI_FLAG_FOURIER_RESIDENT forces the Fourier-resident form for any 1-D
 * ETDRK2 plan (complex L included: the Fourier-space rollout of
 * docs/examples/performance_hints.ipynb cells 10-16, which keeps imaginary
 * Nyquist parts that a physical round trip drops), and
 * EXMI_FLAG_PHYSICAL_ROLLOUT forces the literal round trip for every
 * strategy. Single steps are unaffected. */
#define EXMI_FLAG_FOURIER_RESIDENT 2
#define EXMI_FLAG_PHYSICAL_ROLLOUT 4

/*
 * Stepper description: the constructor arguments of BaseStepper
 * (_base_stepper.py:27-157) plus the nonlinear function's own arguments and
 * the batch. The linear operator is given as host float64 tables of the
 * diagonal L: its real part (host_linear_operator) and, for odd derivative
 * terms, its imaginary part (host_linear_operator_imag), each of shape
 * (lin_batch, lin_channels, Nr), Nr = N^(D-1) * (N/2+1). lin_batch is 1
 * (shared) or `batch` (per-trajectory coefficients, the eqx.filter_vmap case).
 */
typedef struct exmi_plan_desc {
  int32_t num_spatial_dims;   /* D: 1, 2 or 3 (3: ETDRK0-4, N 16..256)       */
  int32_t num_points;         /* N: any size 2..4096 in every dimension.
                                 The fused kernels cover 1-D powers of two
                                 16..4096, 48 and 50, 2-D powers of two
                                 32..512 and 3-D powers of two 16..256; other
                                 sizes run the multi-pass plan with the
                                 runtime-size mixed-radix transforms (any
                                 factorisation, e.g. N = 100, 513 or an odd
                                 prime), per-step launches, the multi-pass
                                 adjoint, no Fourier-resident rollout      */
  int32_t num_channels;       /* C                                           */
  int32_t batch;              /* B                                           */
  int32_t order;              /* ETDRK order 0..4                            */
  int32_t nonlin_kind;        /* enum exmi_nonlin_kind                       */
  int32_t conv_single_channel;
  int32_t conv_conservative;
  int32_t zero_mode_fix;
  int32_t num_poly;           /* PolynomialNonlinearFun coefficient count    */
  int32_t injection_mode;     /* Kolmogorov forcing wavenumber k_f           */
  int32_t lin_batch;          /* 1 or batch                                  */
  int32_t lin_channels;       /* 1 or num_channels                           */
  int32_t num_circle_points;  /* contour points (reference default 16)       */
  int32_t flags;              /* EXMI_FLAG_*                                 */
  double domain_extent;       /* L                                           */
  double dt;
  double circle_radius;       /* contour radius (reference default 1.0)      */
  double dealiasing_fraction; /* < 0: no dealiasing (reference None)         */
  double scale;               /* convection / gradient-norm / vorticity b    */
  double general_scales[3];   /* GeneralNonlinearFun scale_list              */
  double poly[EXMI_MAX_POLY]; /* PolynomialNonlinearFun coefficients         */
  double injection_scale;     /* Kolmogorov forcing amplitude                */
  const double* host_linear_operator; /* (lin_batch, lin_channels, Nr) f64  */
  /* NULL for a real L; else the imaginary parts of L, same shape. A complex L
   * (odd derivatives: advection, dispersion; stepper/_advection.py:88-95,
   * _dispersion.py:102-117, _korteweg_de_vries.py) is built for ETDRK orders
   * 0-4: 1-D plans with N <= 1024 (EXMI_FLAG_FOURIER_RESIDENT: ETDRK2), 2-D
   * and 3-D plans. exp(L dt) (and exp(L dt / 2) for orders 3-4)
   * is then complex while the phi-coefficients stay real: the reference keeps
   * the real part of every contour sample (etdrk/_etdrk_1.py:66-76,
   * _etdrk_2.py:71-89). The table layout gains Im E after the real
   * coefficients (exmi_plan_num_coefs counts it). The 1-D adjoint of the
   * linear, convection and general plans multiplies by conj(exp(L dt))
   * (exmi_step_vjp). */
  const double* host_linear_operator_imag;
  double reaction_rates[2];   /* Gray-Scott feed and kill rates              */
  /* NULL, or (batch,) host float64 factors f_b: trajectory b steps with the
   * nonlinear term f_b N(u) (the plan's nonlinear parameters times f_b, e.g.
   * per-trajectory convection or gradient-norm scales in an ensemble,
   * eqx.filter_vmap over them). Folded into the coefficients that multiply
   * N (every ETDRK stage is linear in N); needs lin_batch == batch.
   * exmi_nonlinear returns the unscaled N. */
  const double* host_nonlin_factor;
} exmi_plan_desc;

typedef struct exmi_plan exmi_plan;

/* Build a plan: validates the description, uploads L and forms exp(L dt)
 * and the ETDRK phi-coefficients by the contour-integral mean in float64 on
 * the device, stored as float32 tables.
 * Replaces BaseStepper.__init__ (_base_stepper.py:27-157), BaseETDRK.__init__
 * (etdrk/_base_etdrk.py:18-62) and ETDRK{1,2,3,4}.__init__ (etdrk/_etdrk_1.py:13-76,
 * _etdrk_2.py:13-89, _etdrk_3.py:13-189, _etdrk_4.py:13-196). Synchronous. */
int exmi_plan_create(const exmi_plan_desc* desc, exmi_plan** out);

/* Release a plan and its device tables, workspace and cached graphs.
 * (The reference's pytrees are garbage collected; no counterpart.) */
int exmi_plan_destroy(exmi_plan* plan);

/* One step on B physical states: u_out = step(u_in), (B, C, N^D) float32.
 * in/out trajectory strides are in floats (C*N^D for contiguous batches);
 * u_in may equal u_out.
 * Replaces BaseStepper.__call__/step (_base_stepper.py:201-220, 241-271)
 * under jax.vmap. */
int exmi_step(exmi_plan* plan, const float* u_in, int64_t in_bstride, float* u_out,
              int64_t out_bstride, void* stream);

/* One step in Fourier space on (B, C, Nr) complex64 spectra (float pairs);
 * strides in complex elements (2-D: contiguous spectra). The step runs on
 * the spectra with no outer transforms: every mode is carried, imaginary
 * parts of the DC and Nyquist bins included, as the reference's
 * step_fourier carries them; uh_in may equal uh_out. Coupled-reaction 1-D
 * plans return EXMI_ERR_UNSUPPORTED.
 * Replaces BaseStepper.step_fourier (_base_stepper.py:222-239) /
 * BaseETDRK.step_fourier (etdrk/_base_etdrk.py:64-80). */
int exmi_step_fourier(exmi_plan* plan, const float* uh_in, int64_t in_bstride, float* uh_out,
                      int64_t out_bstride, void* stream);

/* Dealiased nonlinear term N(u_hat) on (B, C, Nr) complex64 spectra (2-D:
 * contiguous spectra, real L; coupled-reaction 1-D plans return
 * EXMI_ERR_UNSUPPORTED).
 * Replaces BaseNonlinearFun.__call__ (nonlin_fun/_base.py:139-159) of the
 * plan's nonlinear function. */
int exmi_nonlinear(exmi_plan* plan, const float* uh_in, int64_t in_bstride, float* nh_out,
                   int64_t out_bstride, void* stream);

/* Trajectory: traj[b, t] for t = 0..n_steps-1 (+1 with include_init),
 * traj float32 (B, n_steps[+include_init], C, N^D) contiguous; u0 (B, C, N^D)
 * with stride u0_bstride floats. With use_graph != 0 the n_steps launches are
 * captured once in a hipGraph (cached per plan for the same pointers and n)
 * and replayed on `stream`.
 * Replaces exponax.rollout (_utils.py:92-186, takes_aux=False). */
int exmi_rollout(exmi_plan* plan, const float* u0, int64_t u0_bstride, float* traj, int32_t n_steps,
                 int32_t include_init, int32_t use_graph, void* stream);

/* exmi_rollout (no graph) with failure detection: first_bad (B,) device
 * int32 receives, per trajectory, the first time index t of traj whose
 * state holds a value that is not finite or exceeds `bound` in magnitude
 * (bound <= 0: non-finite values only), or -1; with include_init the given
 * initial state (t = 0) is not checked. The multi-step 1-D kernels check
 * every state as they write it (no extra pass); other strategies and
 * dimensions scan the written states once (exmi_first_bad_step).
 * (No reference counterpart: failure detection for long batched rollouts,
 * in which trajectories blow up one by one.) */
int exmi_rollout_checked(exmi_plan* plan, const float* u0, int64_t u0_bstride, float* traj, int32_t n_steps,
                         int32_t include_init, float bound, int32_t* first_bad, void* stream);

/* Capture and instantiate the rollout graph for these arguments without
 * launching it, so that a later exmi_rollout(..., use_graph=1) with the same
 * arguments only replays. Synchronous on the host; enqueues no GPU work.
 * (No reference counterpart: jax.jit/lax.scan compile on first call.) */
int exmi_rollout_prepare(exmi_plan* plan, const float* u0, int64_t u0_bstride, float* traj, int32_t n_steps,
                         int32_t include_init);

/* Rollouts of several plans advanced together, e.g. instances of different
 * grid sizes in one batch: plan i rolls u0s[i] (stride u0_bstrides[i]) into
 * trajs[i] exactly as exmi_rollout would. use_graph: 0 = every launch in
 * order on `stream`; 1 = one captured hipGraph in which each plan's launches
 * form a parallel branch (cached by the first plan for the same arguments);
 * 2 = one cached graph per plan, the graphs launched concurrently on the
 * plans' own streams (forked from and joined into `stream`); 3 = as 2 with
 * direct launches instead of graphs. A plan may appear once.
 * (The reference runs each grid size separately: arrays of different shapes
 * cannot share one jax.vmap, docs/examples/performance_hints.ipynb cell 22.) */
int exmi_rollout_group(exmi_plan* const* plans, int32_t n_plans, const float* const* u0s,
                       const int64_t* u0_bstrides, float* const* trajs, int32_t n_steps, int32_t include_init,
                       int32_t use_graph, void* stream);

/* Final state after n_steps: u_out (B, C, N^D) contiguous.
 * Replaces exponax.repeat (_utils.py:189-254, takes_aux=False). */
int exmi_repeat(exmi_plan* plan, const float* u0, int64_t u0_bstride, float* u_out, int32_t n_steps,
                int32_t use_graph, void* stream);

/* Adjoint of one step (1-D: also the polynomial reactions): g_in =
 * (d step(u) / d u)^T g_out at u = u_in, all
 * physical (B, C, N^D) float32 with batch strides in floats; g_in may equal
 * g_out (not u_in). Inner products are the plain sums over grid points, as
 * jax.grad of a loss sum(step(u) * g_out) would give.
 * Built for ETDRK orders 0-4 in 1-D, 2-D and 3-D with no nonlinearity,
 * every convection form, the gradient norm, polynomials, the general form,
 * the 2-D vorticity (Kolmogorov) convection and the projected 3-D
 * (Navier-Stokes velocity) convection, real or complex L (the
 * transpose multiplies by conj(exp(L dt))): fused kernels for 1-D orders
 * 0-2 and 2-D gradient-norm ETDRK2, the multi-pass transpose otherwise (for
 * a fused plan through a multi-pass twin of the same stepper, built
 * synchronously on the first call), Cahn-Hilliard and the coupled
 * reactions included.
 * Replaces jax.vjp of BaseStepper.__call__ (_base_stepper.py:241-271), i.e.
 * the transposes JAX derives for etdrk/_etdrk_2.py:91-102 and the nonlinear
 * functions (nonlin_fun/_convection.py, _gradient_norm.py). */
int exmi_step_vjp(exmi_plan* plan, const float* u_in, int64_t in_bstride, const float* g_out, int64_t g_bstride,
                  float* g_in, int64_t gin_bstride, void* stream);

/* Gradient of a loss through a rollout with segment checkpointing:
 * g_u0 = sum_t (d u_t / d u0)^T g_t, where u_t are the states that
 * exmi_rollout(plan, u0, ..., n_steps, include_init) produces. g_traj holds
 * the cotangents of that trajectory, (B, n_steps[+1], C, N^D) contiguous, or
 * is NULL. g_final, (B, C, N^D) contiguous or NULL, is an extra cotangent of
 * the final state (exponax.repeat). g_u0 is (B, C, N^D) contiguous.
 * Every checkpoint_every-th state is kept (<= 0: all states, i.e. one
 * segment). Each segment is recomputed from its checkpoint on the way back.
 * Workspace: (ceil(n/K) + K + 1) * B * C * N^D floats, kept by the plan.
 * Replaces jax.grad through exponax.rollout / repeat (_utils.py:92-254,
 * lax.scan at :149-172, :239-251). */
int exmi_rollout_vjp(exmi_plan* plan, const float* u0, int64_t u0_bstride, int32_t n_steps, int32_t include_init,
                     const float* g_traj, const float* g_final, float* g_u0, int32_t checkpoint_every, void* stream);

/* Batched real FFTs with the reference conventions: forward rfftn over the
 * last D axes, norm="backward"; inverse irfftn with 1/N^D scaling, Im of the
 * last axis' DC/Nyquist ignored. `count` transforms, contiguous: real
 * (count, N^D), spectra (count, N^(D-1), N/2+1) in the reference's axis
 * order. D = 3: powers of two 16..256 (the inverse through a stream-ordered
 * copy of the spectra). D = 1: powers of two 16..4096, 48 and 50. D = 2: powers of two
 * 32..512, count <= 65535, through a stream-ordered (hipMallocAsync)
 * intermediate of the spectra's size. Every other N in 2..4096 (any D) runs
 * the runtime-size mixed-radix transforms (the inverse of D >= 2 through a
 * stream-ordered copy of the spectra).
 * Replace exponax.fft / exponax.ifft (_spectral.py:614-656, 659-721). */
int exmi_rfft(int32_t num_spatial_dims, int32_t num_points, int32_t count, const float* x, float* x_hat,
              void* stream);
int exmi_irfft(int32_t num_spatial_dims, int32_t num_points, int32_t count, const float* x_hat, float* x,
               void* stream);

/* exmi_irfft that allocates nothing, for callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). It has no counterpart in the reference. Same contract and
 * results as exmi_irfft; the 2-D [ky][x] intermediate and the copy of the
 * 3-D spectra (the line transforms run in place) live in a library-owned
 * device array of EXMI_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, and a request needing more fails with EXMI_ERR_INVALID before any
 * work is enqueued (2-D: (N/2+1) N complex64 per field, so 31 fields of
 * 512^2; 3-D: N^2 (N/2+1) complex64 per field, so 3 fields of 128^3 and
 * none of 256^3). Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race. The first call on
 * a device runs a synchronous set-up (symbol addresses, a host-to-device
 * copy of the twiddle tables); make it outside graph capture, e.g. with
 * count = 0. */
#define EXMI_NOALLOC_SCRATCH_BYTES (32u << 20)
int exmi_irfft_noalloc(int32_t num_spatial_dims, int32_t num_points, int32_t count, const float* x_hat, float* x,
                       void* stream);

/* Normalized RMSE of B state pairs: out[b] = sum_c ||pred_bc - ref_bc||_2 /
 * ||ref_bc||_2 over the n = N^D points of each channel (the (L/N)^D factor
 * cancels), float64 accumulation, float32 result. pred/ref (B, C, n) with
 * batch strides in floats; out (B,) device floats.
 * Replaces exponax.metrics.nRMSE (metrics/_spatial.py:537-586) under
 * jax.vmap. */
int exmi_nrmse(int32_t batch, int32_t channels, int64_t n, const float* pred, int64_t pred_bstride,
               const float* ref, int64_t ref_bstride, float* out, void* stream);

/* Poisson solve in Fourier space, in place on `count` contiguous
 * (N^(D-1), N/2+1) complex64 spectra: u_hat = -f_hat / sum_d (i kappa_d)^p
 * (p = order, even), 0 where that operator vanishes (mean-zero solution).
 * Replaces Poisson.step_fourier (_poisson.py:12-80). */
int exmi_poisson_fourier(int32_t num_spatial_dims, int32_t num_points, int32_t count, double domain_extent,
                         int32_t order, float* uh, void* stream);

/* out = x + a y over n floats (device arrays; out may alias x or y).
 * Replaces the forcing of ForcedStepper.step / step_fourier
 * (_forced_stepper.py:33-56): u + dt f. */
int exmi_axpy(int64_t n, float a, const float* x, const float* y, float* out, void* stream);

/* The wave equation's exact Fourier-space step, in place on `batch`
 * contiguous (2, N^(D-1), N/2+1) complex64 spectra of (h, v): the
 * diagonalisation, exp(+- i c |k| dt), the inverse diagonalisation and the
 * mean mode's drift h += dt v.
 * Replaces Wave.step_fourier (stepper/_wave.py:130-197). */
int exmi_wave_step_fourier(int32_t num_spatial_dims, int32_t num_points, int32_t batch, double domain_extent,
                           double speed_of_sound, double dt, float* uh, void* stream);

/* Seeded truncated-Fourier-series initial conditions on the device (the
 * role of ic/_truncated_fourier_series.py:17-100 with the library's own
 * draws): out (count, N^D) = sum_m amp[f, m] cos(2 pi k_m . x / L +
 * phase[f, m]) divided by its max |u| per field f. `modes` is a HOST array
 * (nmodes, D) of distinct half-plane representatives (last non-zero
 * component positive) with every |k_d| < N/2; amp and phase are device
 * arrays (count, nmodes). Sizes as for exmi_irfft. */
int exmi_ic_tfs(int32_t num_spatial_dims, int32_t num_points, int32_t count, int32_t nmodes, const int32_t* modes,
                const float* amp, const float* phase, float* out, void* stream);

/* Spectrum of `count` fields of N^D points: out (count, N/2+1) float32,
 * power = 0.5 |u_hat|^2 / (s_rec s_norm) when `power`, else |u_hat| / s_rec,
 * with the reference's scaling arrays (reconstruction: N at k = 0, N/2 and
 * N/2 elsewhere on the last axis, N on the others; norm compensation:
 * N^D). For D = 2, 3 the modes are summed over radial buckets
 * [b - 1/2, b + 1/2) of the wavenumber norm, b = 0..N/2
 * (radial_binning="sum"; exmi_spectrum_binned with `average`:
 * radial_binning="average", NaN for an empty bucket, as nanmean; D = 1
 * ignores it); sizes as for exmi_rfft.
 * Replaces exponax.get_spectrum (_spectral.py:866-1030). */
int exmi_spectrum_binned(int32_t num_spatial_dims, int32_t num_points, int32_t count, const float* x, float* out,
                         int32_t power, int32_t average, void* stream);
int exmi_spectrum(int32_t num_spatial_dims, int32_t num_points, int32_t count, const float* x, float* out,
                  int32_t power, void* stream);

/* Per trajectory b of a batch of trajectories x (b, t, 0..n-1) with batch
 * and time strides in floats: first_bad[b] (device int32) = the smallest
 * t < n_times whose state holds a value that is not finite or exceeds
 * `bound` in magnitude (bound <= 0: non-finite values only), or -1.
 * (No reference counterpart: the failure detection of exmi_rollout_checked
 * for any trajectory buffer, e.g. a corrected coarse rollout.) */
int exmi_first_bad_step(int32_t batch, int32_t n_times, int64_t n, const float* x, int64_t bstride, int64_t tstride,
                        float bound, int32_t* first_bad, void* stream);

/* Spectral up/down-sampling of `count` fields from n_old to n_new points
 * per axis (powers of two): the modes of the reference's blocks
 * (_spectral.py:530-611; 1-D: k <= min(n_old, n_new)/2) are carried over
 * with the norm compensation (n_new/n_old)^D; with oddball_zero the Nyquist
 * modes of the even source (upsampling) or target (downsampling) grid are
 * zeroed. x (count, n_old^D) -> y (count, n_new^D), contiguous. D = 1:
 * 16..4096, D = 2: 32..512, D = 3: 16..256.
 * Replaces exponax.map_between_resolutions (_interpolation.py:195-285). */
int exmi_map_resolution(int32_t num_spatial_dims, int32_t n_old, int32_t n_new, int32_t count, const float* x,
                        float* y, int32_t oddball_zero, void* stream);

/* ---- Learned corrector (no reference counterpart: the reference's closest
 * analogue is the periodic Conv1d emulator of
 * docs/examples/learning_burgers_autoregressive_neural_operator.ipynb).
 * A periodic 1-D CNN y = x + f(x), 5 Conv1d layers of kernel 5 with widths
 * 1 -> 32 -> 32 -> 32 -> 32 -> 1 and GELU (tanh form) after layers 1-4, on
 * num_points in {32, 64, 128}. Weights and activations enter the
 * convolutions as bfloat16, accumulation is float32; the 32 -> 32 layers run
 * on the matrix cores. params: EXMI_CNN_NUM_PARAMS float32 values,
 * W1 (32,1,5) b1 (32) W2 (32,32,5) b2 W3 b3 W4 b4 W5 (1,32,5) b5 (1), the
 * PyTorch Conv1d layout. Samples are (count, num_points) contiguous. */
#define EXMI_CNN_NUM_PARAMS 15809

/* Bytes of the training workspace for `count` samples. */
int64_t exmi_cnn_workspace_bytes(int32_t num_points, int32_t count);

/* y = x + f(x). With a workspace (exmi_cnn_workspace_bytes) the
 * intermediates that exmi_cnn_backward needs are stored in it. */
int exmi_cnn_forward(const float* params, int32_t num_points, int32_t count, const float* x, float* y,
                     void* workspace, void* stream);

/* Gradients of sum(y * dy) for the samples of the last exmi_cnn_forward into
 * `workspace`: dparams (EXMI_CNN_NUM_PARAMS, overwritten, summed over the
 * samples) and, when dx is not NULL, dx (count, num_points). The weight
 * gradients are reduced over per-wavefront partial sums in a fixed order
 * (deterministic). */
int exmi_cnn_backward(const float* params, int32_t num_points, int32_t count, const float* dy, float* dx,
                      float* dparams, void* workspace, void* stream);

/* Corrected coarse rollout: traj[b, t] = cnn(step(traj[b, t-1])) for
 * t = 0..n_steps-1 (+1 with include_init), starting from u0, with `plan` a
 * 1-D single-channel plan of 32, 64 or 128 points and `params` the
 * network's EXMI_CNN_NUM_PARAMS values. traj (B, n_steps[+1], 1, N)
 * contiguous. Two launches per step: the plan's fused step and the network.
 * (No reference counterpart; the rollout semantics are exponax.rollout's,
 * _utils.py:92-186.) */
int exmi_corrected_rollout(exmi_plan* plan, const float* params, const float* u0, int64_t u0_bstride, float* traj,
                           int32_t n_steps, int32_t include_init, void* stream);

/* Copy the plan's float32 operator tables to host memory:
 * (lin_batch, lin_channels, ncoef, Nr) with ncoef = exmi_plan_num_coefs().
 * Coefficient order: ETDRK0 [E]; ETDRK1 [E, c1]; ETDRK2 [E, c1, c2];
 * ETDRK3 [E, E_half, c1..c5]; ETDRK4 [E, E_half, c1, c4, c5, c6].
 * Exposes BaseETDRK._exp_term / ETDRK*._coef_* (etdrk/_etdrk_2.py:88-89 etc.)
 * for inspection. Synchronous. */
int exmi_plan_num_coefs(const exmi_plan* plan);
int exmi_plan_get_tables(const exmi_plan* plan, float* host_out, int64_t host_count);

/* Rollout strategy of a plan: -1 = one launch per step (each a fused step
 * kernel; the per-step physical round trip), 1 or 2 = one multi-step launch
 * keeping 1 or 2 transforms per lane group on chip for all steps (the
 * Fourier-resident form where the plan has it, see
 * EXMI_FLAG_PHYSICAL_ROLLOUT). 2-D: 1 = the workgroup-resident kernel, 2 =
 * the spectrum-carrying launch chain, 3 = the same chain as one cooperative
 * launch (single-channel gradient norm and vorticity, N = 128 and 256, when
 * the whole grid fits the device; inside exmi_rollout_group with several
 * plans it runs as 2). 3-D and runtime-size (multi-pass) plans: 2 = the
 * spectra carried from step to step, each emitted state inverse-transformed
 * (real L; for an even N the Nyquist modes dealiased). 0 = not chosen
 * yet (the first exmi_rollout / exmi_rollout_prepare times the candidates on
 * a probe of min(n_steps, 16) steps and keeps the fastest). Setting a
 * strategy drops a cached graph. With the environment variable
 * EXMI_TUNE_CACHE naming a file, choices persist across processes: the
 * first rollout of a plan looks its key up (GPU architecture, D, N, C, batch,
 * kernel kind, order, per-trajectory tables, Fourier-resident) before
 * probing, and appends a new choice as a "<key> <strategy>" line.
 * (No reference counterpart: tuning knobs of this implementation.) */
int exmi_plan_rollout_strategy(const exmi_plan* plan);
int exmi_plan_set_rollout_strategy(exmi_plan* plan, int32_t strategy);

/* Transform knobs of a plan on the runtime-size transforms (grid sizes
 * outside the fused kernels' set): the Stockham radices (passes of at most
 * max_radix points; a larger prime factor is a direct DFT pass of its own)
 * and the lines one workgroup holds in LDS. exmi_plan_create times a set of
 * candidates on the plan's batch and keeps the fastest (with
 * EXMI_TUNE_CACHE set, it looks the choice up first and stores a new one,
 * keyed by GPU architecture, D, N and the number of fields).
 * exmi_plan_fft_config writes [code, source, lines per workgroup, passes,
 * radix_1 .. radix_passes] into info (up to len entries) and returns the
 * entry count, 0 for plans on the fused kernels; code = reg * 100000000 +
 * max_radix * 100000 + points per workgroup (reg 1: passes of radix 9..17
 * and 19 as register butterflies, in a kernel of lower occupancy; reg 0:
 * such passes as direct sums); source 0 default, 1 timed, 2 tuning cache,
 * 3 set.
 * exmi_plan_set_fft_config applies a code (drops a cached graph).
 * (No reference counterpart: XLA picks the reference's FFT plans.) */
int exmi_plan_fft_config(const exmi_plan* plan, int32_t* info, int32_t len);
int exmi_plan_set_fft_config(exmi_plan* plan, int32_t code);

/* Message for the last error on this thread ("" if none). */
const char* exmi_last_error(void);

/* Library / device information: writes a short description into buf. */
int exmi_info(char* buf, int32_t buflen);

#ifdef __cplusplus
}
#endif

#endif /* EXMI_H */
