// Instantiates the fused 1-D step and adjoint kernels and the standalone batched real
// FFT kernels for one transform size, selected at compile time with
// -DEXMI_M=<N/2> -DEXMI_PT=<points per lane>. The build compiles this file
// once per supported size.
#include "adjoint1d.hpp"
#include "step1d.hpp"

#ifndef EXMI_M
#error "EXMI_M must be defined"
#endif

namespace exmi {

using StepFn = hipError_t (*)(const Step1DArgs&, hipStream_t);

// Batched real FFTs of count rows of 2M points (forward: to M + 1 complex,
// inverse: from them). kcut: the inverse reads the entries above kcut as
// zero and the forward does not store them (the planes above a dealias
// cutoff; kcut >= M: all). rpg > 0: the inverse's output rows come in
// groups of rpg rows, group q starting at out + q * gstride floats (a
// strided batch of states); rpg = 0: contiguous rows.
template <int M, int PT, bool INV>
__global__ void __launch_bounds__((M / PT) > 64 ? (M / PT) : 64)
    fft1d_kernel(const float* __restrict__ in, float* __restrict__ out, int count, RealTables rt, int kcut,
                 long long rpg, long long gstride) {
  constexpr int T = M / PT;
  constexpr int WG = T > 64 ? T : 64;
  constexpr int G = WG / T;
  constexpr int LDSN = LdsSize<M>::value;
  __shared__ cf lds_all[G * LDSN];
  const int g = threadIdx.x / T, j = threadIdx.x % T;
  const int t_raw = blockIdx.x * G + g;
  const bool active = t_raw < count;
  const int t = active ? t_raw : 0;
  cf* lds = lds_all + g * LDSN;
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> tr;
  tr.load(j, rt);
  cf v[1][PT];
  cf nyq[1] = {cf{0.f, 0.f}};
  if (!INV) {
    const float2* x = reinterpret_cast<const float2*>(in + (long long)t * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      float2 q = x[j + s * T];
      v[0][s] = cf{q.x, q.y};
    }
    rfft<M, PT, 1>(v, nyq, ln, lds, tr);  // 2 X
    if (active) {
      cf* y = reinterpret_cast<cf*>(out) + (long long)t * (M + 1);
#pragma unroll
      for (int s = 0; s < PT; ++s)
        if (j + s * T <= kcut) y[j + s * T] = 0.5f * v[0][s];
      if (j == 0 && M <= kcut) y[M] = 0.5f * nyq[0];
    }
  } else {
    const cf* x = reinterpret_cast<const cf*>(in) + (long long)t * (M + 1);
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = j + s * T <= kcut ? x[j + s * T] : cf{0.f, 0.f};
    nyq[0] = M <= kcut ? x[M] : cf{0.f, 0.f};
    irfft<M, PT, 1>(v, nyq, ln, lds, tr);  // N x
    if (active) {
      constexpr float rn = 1.f / (2.f * M);
      const long long ro = rpg > 0 ? (t / rpg) * gstride + (t % rpg) * 2 * M : (long long)t * 2 * M;
      float2* y = reinterpret_cast<float2*>(out + ro);
#pragma unroll
      for (int s = 0; s < PT; ++s) y[j + s * T] = make_float2(v[0][s].re * rn, v[0][s].im * rn);
    }
  }
}

#define EXMI_CAT2(a, b) a##b
#define EXMI_CAT(a, b) EXMI_CAT2(a, b)

template <int M, int PT, int KIND>
static StepFn pick_order(int order) {
  switch (order) {
    case 1: return &launch_step1d<M, PT, KIND, 1>;
    case 2: return &launch_step1d<M, PT, KIND, 2>;
    case 3: return &launch_step1d<M, PT, KIND, 3>;
    case 4: return &launch_step1d<M, PT, KIND, 4>;
    default: return nullptr;
  }
}

StepFn EXMI_CAT(get_step1d_, EXMI_M)(int kind, int order) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  if (order == 0) return &launch_step1d<M, PT, NL_ZERO, 0>;
  switch (kind) {
    case NL_ZERO: return pick_order<M, PT, NL_ZERO>(order);
    case NL_CONV_CONS: return pick_order<M, PT, NL_CONV_CONS>(order);
    case NL_CONV_NONCONS: return pick_order<M, PT, NL_CONV_NONCONS>(order);
    case NL_GRADNORM: return pick_order<M, PT, NL_GRADNORM>(order);
    case NL_POLY: return pick_order<M, PT, NL_POLY>(order);
    case NL_GENERAL: return pick_order<M, PT, NL_GENERAL>(order);
    default: return nullptr;
  }
}

// Nonlinear-only evaluation shares the ORDER-1 instantiation's register plan;
// it is dispatched through the same kernel with nonlin_only = 1.

using RollFn = hipError_t (*)(const Roll1DArgs&, hipStream_t);

template <int M, int PT, int KIND, int NT>
static RollFn pick_roll_order(int order) {
  switch (order) {
    case 1: return &launch_roll1d<M, PT, KIND, 1, NT>;
    case 2: return &launch_roll1d<M, PT, KIND, 2, NT>;
    case 3: return &launch_roll1d<M, PT, KIND, 3, NT>;
    case 4: return &launch_roll1d<M, PT, KIND, 4, NT>;
    default: return nullptr;
  }
}

template <int M, int PT, int NT>
static RollFn pick_roll(int kind, int order) {
  if (order == 0) return &launch_roll1d<M, PT, NL_ZERO, 0, NT>;
  switch (kind) {
    case NL_ZERO: return pick_roll_order<M, PT, NL_ZERO, NT>(order);
    case NL_CONV_CONS: return pick_roll_order<M, PT, NL_CONV_CONS, NT>(order);
    case NL_CONV_NONCONS: return pick_roll_order<M, PT, NL_CONV_NONCONS, NT>(order);
    case NL_GRADNORM: return pick_roll_order<M, PT, NL_GRADNORM, NT>(order);
    case NL_POLY: return pick_roll_order<M, PT, NL_POLY, NT>(order);
    case NL_GENERAL: return pick_roll_order<M, PT, NL_GENERAL, NT>(order);
    default: return nullptr;
  }
}

template <int M, int PT, int NT>
static RollFn pick_roll_fr(int kind) {
  switch (kind) {
    case NL_ZERO: return &launch_roll1d<M, PT, NL_ZERO, 2, NT, true>;
    case NL_CONV_CONS: return &launch_roll1d<M, PT, NL_CONV_CONS, 2, NT, true>;
    case NL_CONV_NONCONS: return &launch_roll1d<M, PT, NL_CONV_NONCONS, 2, NT, true>;
    case NL_GRADNORM: return &launch_roll1d<M, PT, NL_GRADNORM, 2, NT, true>;
    case NL_POLY: return &launch_roll1d<M, PT, NL_POLY, 2, NT, true>;
    case NL_GENERAL: return &launch_roll1d<M, PT, NL_GENERAL, 2, NT, true>;
    default: return nullptr;
  }
}

// ETDRK2 rollouts of the common nonlinear kinds with the default 2/3
// dealias cutoff as a compile-time constant.
template <int M, int PT, int NT, bool FR>
static RollFn pick_roll_kc(int kind) {
  constexpr int KC = default_kmax(M);
  switch (kind) {
    case NL_CONV_CONS: return &launch_roll1d<M, PT, NL_CONV_CONS, 2, NT, FR, KC>;
    case NL_CONV_NONCONS: return &launch_roll1d<M, PT, NL_CONV_NONCONS, 2, NT, FR, KC>;
    case NL_GRADNORM: return &launch_roll1d<M, PT, NL_GRADNORM, 2, NT, FR, KC>;
    default: return nullptr;
  }
}

// Multi-step kernels carrying 1 or 2 transforms per lane group; `fr`
// selects the Fourier-resident variant (ETDRK2 only); `std_kc` says that the
// plan's dealias cutoff is default_kmax(M).
RollFn EXMI_CAT(get_roll1d_, EXMI_M)(int kind, int order, int nt, int fr, int std_kc) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  if (std_kc && order == 2) {
    RollFn f = nullptr;
    if (nt == 1) f = fr ? pick_roll_kc<M, PT, 1, true>(kind) : pick_roll_kc<M, PT, 1, false>(kind);
    if constexpr (M / PT <= 64) {
      if (nt == 2) f = fr ? pick_roll_kc<M, PT, 2, true>(kind) : pick_roll_kc<M, PT, 2, false>(kind);
    }
    if (f) return f;
  }
  if (fr) {
    if (order != 2) return nullptr;
    if (nt == 1) return pick_roll_fr<M, PT, 1>(kind);
    if constexpr (M / PT <= 64) {
      if (nt == 2) return pick_roll_fr<M, PT, 2>(kind);
    }
    return nullptr;
  }
  if (nt == 1) return pick_roll<M, PT, 1>(kind, order);
  if constexpr (M / PT <= 64) {
    if (nt == 2) return pick_roll<M, PT, 2>(kind, order);
  }
  return nullptr;
}

// Complex linear operator (exp(L dt) complex, real phi-coefficients): ETDRK
// orders 0-2, one transform per lane group, sizes up to M = 512.
template <int M, int PT, int KIND>
static StepFn pick_step_c(int order) {
  switch (order) {
    case 1: return &launch_step1d<M, PT, KIND, 1, EXMI_LAYOUT_1D, true>;
    case 2: return &launch_step1d<M, PT, KIND, 2, EXMI_LAYOUT_1D, true>;
    case 3: return &launch_step1d<M, PT, KIND, 3, EXMI_LAYOUT_1D, true>;
    case 4: return &launch_step1d<M, PT, KIND, 4, EXMI_LAYOUT_1D, true>;
    default: return nullptr;
  }
}
template <int M, int PT, int KIND>
static RollFn pick_roll_c(int order) {
  switch (order) {
    case 1: return &launch_roll1d<M, PT, KIND, 1, 1, false, -1, EXMI_LAYOUT_1D, true>;
    case 2: return &launch_roll1d<M, PT, KIND, 2, 1, false, -1, EXMI_LAYOUT_1D, true>;
    case 3: return &launch_roll1d<M, PT, KIND, 3, 1, false, -1, EXMI_LAYOUT_1D, true>;
    case 4: return &launch_roll1d<M, PT, KIND, 4, 1, false, -1, EXMI_LAYOUT_1D, true>;
    default: return nullptr;
  }
}

StepFn EXMI_CAT(get_step1d_c_, EXMI_M)(int kind, int order) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  if constexpr (M <= 512) {
    if (order == 0) return &launch_step1d<M, PT, NL_ZERO, 0, EXMI_LAYOUT_1D, true>;
    switch (kind) {
      case NL_ZERO: return pick_step_c<M, PT, NL_ZERO>(order);
      case NL_CONV_CONS: return pick_step_c<M, PT, NL_CONV_CONS>(order);
      case NL_CONV_NONCONS: return pick_step_c<M, PT, NL_CONV_NONCONS>(order);
      case NL_GRADNORM: return pick_step_c<M, PT, NL_GRADNORM>(order);
      case NL_POLY: return pick_step_c<M, PT, NL_POLY>(order);
      case NL_GENERAL: return pick_step_c<M, PT, NL_GENERAL>(order);
      default: return nullptr;
    }
  }
  return nullptr;
}

// Channel-coupled reactions (ETDRK orders 1-4): one trajectory's channels
// are the transforms of one lane group (NT = channels).
template <int M, int PT, int KIND, int NT>
static RollFn pick_roll_r(int order) {
  switch (order) {
    case 1: return &launch_roll1d<M, PT, KIND, 1, NT>;
    case 2: return &launch_roll1d<M, PT, KIND, 2, NT>;
    case 3: return &launch_roll1d<M, PT, KIND, 3, NT>;
    case 4: return &launch_roll1d<M, PT, KIND, 4, NT>;
    default: return nullptr;
  }
}

RollFn EXMI_CAT(get_roll1d_r_, EXMI_M)(int kind, int order) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  if constexpr (M <= 512) {
    if (kind == NL_GRAY_SCOTT) return pick_roll_r<M, PT, NL_GRAY_SCOTT, 2>(order);
    if (kind == NL_BZ) return pick_roll_r<M, PT, NL_BZ, 3>(order);
  }
  return nullptr;
}

// complex L; `fr`: the Fourier-resident ETDRK2 rollout
template <int M, int PT, int KIND>
static RollFn pick_roll_cf() {
  return &launch_roll1d<M, PT, KIND, 2, 1, true, -1, EXMI_LAYOUT_1D, true>;
}

RollFn EXMI_CAT(get_roll1d_c_, EXMI_M)(int kind, int order, int fr) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  if constexpr (M <= 512) {
    if (fr) {
      if (order != 2) return nullptr;
      switch (kind) {
        case NL_ZERO: return pick_roll_cf<M, PT, NL_ZERO>();
        case NL_CONV_CONS: return pick_roll_cf<M, PT, NL_CONV_CONS>();
        case NL_CONV_NONCONS: return pick_roll_cf<M, PT, NL_CONV_NONCONS>();
        case NL_GRADNORM: return pick_roll_cf<M, PT, NL_GRADNORM>();
        case NL_POLY: return pick_roll_cf<M, PT, NL_POLY>();
        case NL_GENERAL: return pick_roll_cf<M, PT, NL_GENERAL>();
        default: return nullptr;
      }
    }
    if (order == 0) return &launch_roll1d<M, PT, NL_ZERO, 0, 1, false, -1, EXMI_LAYOUT_1D, true>;
    switch (kind) {
      case NL_ZERO: return pick_roll_c<M, PT, NL_ZERO>(order);
      case NL_CONV_CONS: return pick_roll_c<M, PT, NL_CONV_CONS>(order);
      case NL_CONV_NONCONS: return pick_roll_c<M, PT, NL_CONV_NONCONS>(order);
      case NL_GRADNORM: return pick_roll_c<M, PT, NL_GRADNORM>(order);
      case NL_POLY: return pick_roll_c<M, PT, NL_POLY>(order);
      case NL_GENERAL: return pick_roll_c<M, PT, NL_GENERAL>(order);
      default: return nullptr;
    }
  }
  return nullptr;
}

using VjpFn = hipError_t (*)(const Vjp1DArgs&, hipStream_t);

template <int M, int PT, int KIND, bool CPLX = false>
static VjpFn pick_vjp_order(int order) {
  switch (order) {
    case 1: return &launch_vjp1d<M, PT, KIND, 1, CPLX>;
    case 2: return &launch_vjp1d<M, PT, KIND, 2, CPLX>;
    default: return nullptr;
  }
}

// Adjoint step kernels: ETDRK orders 0-2, convection (both forms), gradient
// norm, polynomials and the general form; with a complex L (N <= 1024):
// linear, convection and the general form.
VjpFn EXMI_CAT(get_vjp1d_, EXMI_M)(int kind, int order, int cplx) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  if (cplx) {
    if constexpr (M <= 512) {
      if (order == 0) return &launch_vjp1d<M, PT, NL_ZERO, 0, true>;
      switch (kind) {
        case NL_ZERO: return pick_vjp_order<M, PT, NL_ZERO, true>(order);
        case NL_CONV_CONS: return pick_vjp_order<M, PT, NL_CONV_CONS, true>(order);
        case NL_CONV_NONCONS: return pick_vjp_order<M, PT, NL_CONV_NONCONS, true>(order);
        case NL_GENERAL: return pick_vjp_order<M, PT, NL_GENERAL, true>(order);
        default: return nullptr;
      }
    }
    return nullptr;
  }
  if (order == 0) return &launch_vjp1d<M, PT, NL_ZERO, 0>;
  switch (kind) {
    case NL_ZERO: return pick_vjp_order<M, PT, NL_ZERO>(order);
    case NL_CONV_CONS: return pick_vjp_order<M, PT, NL_CONV_CONS>(order);
    case NL_CONV_NONCONS: return pick_vjp_order<M, PT, NL_CONV_NONCONS>(order);
    case NL_GRADNORM: return pick_vjp_order<M, PT, NL_GRADNORM>(order);
    case NL_POLY: return pick_vjp_order<M, PT, NL_POLY>(order);
    case NL_GENERAL: return pick_vjp_order<M, PT, NL_GENERAL>(order);
    default: return nullptr;
  }
}

hipError_t EXMI_CAT(launch_fft1d_, EXMI_M)(bool inverse, const float* in, float* out, int count,
                                           const RealTables& rt, hipStream_t st, int kcut, long long rpg,
                                           long long gstride) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  constexpr int T = M / PT;
  constexpr int WG = T > 64 ? T : 64;
  constexpr int G = WG / T;
  const int grid = (count + G - 1) / G;
  if (inverse)
    hipLaunchKernelGGL((fft1d_kernel<M, PT, true>), dim3(grid), dim3(WG), 0, st, in, out, count, rt, kcut, rpg,
                       gstride);
  else
    hipLaunchKernelGGL((fft1d_kernel<M, PT, false>), dim3(grid), dim3(WG), 0, st, in, out, count, rt, kcut, 0LL,
                       0LL);
  return hipGetLastError();
}

}  // namespace exmi
