// Instantiates the fused 1-D step kernels and the standalone batched real
// FFT kernels for one transform size, selected at compile time with
// -DEXMI_M=<N/2> -DEXMI_PT=<points per lane>. The build compiles this file
// once per supported size.
#include "step1d.hpp"

#ifndef EXMI_M
#error "EXMI_M must be defined"
#endif

namespace exmi {

using StepFn = hipError_t (*)(const Step1DArgs&, hipStream_t);

template <int M, int PT, bool INV>
__global__ void __launch_bounds__((M / PT) > 64 ? (M / PT) : 64)
    fft1d_kernel(const float* __restrict__ in, float* __restrict__ out, int count, float inv_n, const cf* tw,
                 const cf* twr) {
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
  cf v[PT];
  cf nyq = cf{0.f, 0.f};
  if (!INV) {
    const float2* x = reinterpret_cast<const float2*>(in + (long long)t * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      float2 q = x[j + s * T];
      v[s] = cf{q.x, q.y};
    }
    rfft<M, PT>(v, nyq, j, lds, tw, twr);
    if (active) {
      cf* y = reinterpret_cast<cf*>(out) + (long long)t * (M + 1);
#pragma unroll
      for (int s = 0; s < PT; ++s) y[j + s * T] = v[s];
      if (j == 0) y[M] = nyq;
    }
  } else {
    const cf* x = reinterpret_cast<const cf*>(in) + (long long)t * (M + 1);
#pragma unroll
    for (int s = 0; s < PT; ++s) v[s] = x[j + s * T];
    nyq = x[M];
    irfft<M, PT>(v, nyq, j, lds, tw, twr, inv_n);
    if (active) {
      float2* y = reinterpret_cast<float2*>(out + (long long)t * 2 * M);
#pragma unroll
      for (int s = 0; s < PT; ++s) y[j + s * T] = make_float2(v[s].re, v[s].im);
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

hipError_t EXMI_CAT(launch_fft1d_, EXMI_M)(bool inverse, const float* in, float* out, int count, float inv_n,
                                           const cf* tw, const cf* twr, hipStream_t st) {
  constexpr int M = EXMI_M, PT = EXMI_PT;
  constexpr int T = M / PT;
  constexpr int WG = T > 64 ? T : 64;
  constexpr int G = WG / T;
  const int grid = (count + G - 1) / G;
  if (inverse)
    hipLaunchKernelGGL((fft1d_kernel<M, PT, true>), dim3(grid), dim3(WG), 0, st, in, out, count, inv_n, tw, twr);
  else
    hipLaunchKernelGGL((fft1d_kernel<M, PT, false>), dim3(grid), dim3(WG), 0, st, in, out, count, inv_n, tw, twr);
  return hipGetLastError();
}

}  // namespace exmi
