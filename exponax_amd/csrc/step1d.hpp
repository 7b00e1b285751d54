// Fused 1-D ETDRK0-4 pseudo-spectral step for gfx950.
//
// One launch advances every (trajectory, channel) transform of a batch by one
// step (or evaluates the nonlinear term only). A transform is owned by T =
// M/PT lanes and stays in registers for the whole step: the physical state is
// read once (float2 pairs, coalesced), transformed, every nonlinear
// evaluation (dealias, derivative multiply, products, forward transform, post
// dealias) and ETDRK stage combination happens on registers with LDS used
// only for the FFT exchanges, and the new state is written once.
//
// Semantics follow exponax 0.2.0: BaseStepper.step (_base_stepper.py:201-220),
// ETDRK0-4 step_fourier (etdrk/_etdrk_0.py:30-34, _etdrk_1.py:77-82,
// _etdrk_2.py:91-102, _etdrk_3.py:191-212, _etdrk_4.py:198-224) and the
// nonlinear functions of nonlin_fun/ (see nl_eval below).
#pragma once
#include "exmi_internal.hpp"
#include "fft1d.hpp"

namespace exmi {

struct Step1DArgs {
  const void* in;        // float (B,C,N) physical or float2 (B,C,M+1) spectrum
  void* out;             // same conventions
  long long in_bstride;  // elements (float or float2) between trajectories
  long long out_bstride;
  int in_fourier, out_fourier;
  int nonlin_only;       // 1: out = N(in) (Fourier in and out)
  int ntrans;            // B*C
  int channels;          // C
  const float* tables;   // (lin_batch, lin_channels, ncoef, M+1)
  int lin_batch, lin_channels, ncoef;
  const cf* tw;          // exp(-2 pi i m / M), m < M
  const cf* twr;         // exp(-2 pi i k / N), k <= M
  float kscale;          // 2 pi / L (float32, as the reference's derivative operator)
  int kmax;              // dealias: keep |k| <= kmax (kmax >= M: keep all)
  float inv_n;           // 1/N
  float scale;           // b of convection / gradient norm
  float g0, g1, g2;      // general nonlinear scales (b0, b1, b2)
  int zero_mode_fix;
  int npoly;
  float poly[EXMI_MAX_POLY];
};

template <int M, int PT>
struct Xf {
  static constexpr int T = M / PT;
  int j;
  cf* lds;
  const cf* tw;
  const cf* twr;
};

// ----------------------------------------------------------------- helpers
template <int M, int PT>
__device__ __forceinline__ void masked_copy(const cf* S, cf nyq, cf* Z, cf& zn, int j, int kmax) {
  constexpr int T = M / PT;
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int k = j + s * T;
    Z[s] = (k <= kmax) ? S[s] : cf{0.f, 0.f};
  }
  zn = (M <= kmax) ? nyq : cf{0.f, 0.f};
}

// Z = mask * (i kappa) * S
template <int M, int PT>
__device__ __forceinline__ void masked_deriv(const cf* S, cf nyq, cf* Z, cf& zn, int j, int kmax,
                                             float kscale) {
  constexpr int T = M / PT;
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int k = j + s * T;
    const float kap = kscale * (float)k;
    Z[s] = (k <= kmax) ? cf{-kap * S[s].im, kap * S[s].re} : cf{0.f, 0.f};
  }
  const float kn = kscale * (float)M;
  zn = (M <= kmax) ? cf{-kn * nyq.im, kn * nyq.re} : cf{0.f, 0.f};
}

// ------------------------------------------------------ nonlinear functions
// All return the dealiased Fourier-space nonlinear term in (Nn, nn).
template <int KIND, int M, int PT>
__device__ __forceinline__ void nl_eval(const cf* S, cf nyq, cf* Nn, cf& nn, int j, cf* lds,
                                        const cf* __restrict__ tw, const cf* __restrict__ twr,
                                        const Step1DArgs& a) {
  constexpr int T = M / PT;
  if constexpr (KIND == NL_ZERO) {
    // nonlin_fun/_zero.py:33-38
#pragma unroll
    for (int s = 0; s < PT; ++s) Nn[s] = cf{0.f, 0.f};
    nn = cf{0.f, 0.f};
  } else if constexpr (KIND == NL_CONV_CONS) {
    // -b/2 (i kappa) F[u^2]  (nonlin_fun/_convection.py:104-138, :173-202)
    cf Z[PT];
    cf zn;
    masked_copy<M, PT>(S, nyq, Z, zn, j, a.kmax);
    irfft<M, PT>(Z, zn, j, lds, tw, twr, a.inv_n);
#pragma unroll
    for (int s = 0; s < PT; ++s) Z[s] = cf{Z[s].re * Z[s].re, Z[s].im * Z[s].im};
    rfft<M, PT>(Z, zn, j, lds, tw, twr);
    const float c = -a.scale * 0.5f * a.kscale;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const int k = j + s * T;
      const float f = (k <= a.kmax) ? c * (float)k : 0.f;
      Nn[s] = cf{-f * Z[s].im, f * Z[s].re};
    }
    const float fn = (M <= a.kmax) ? c * (float)M : 0.f;
    nn = cf{-fn * zn.im, fn * zn.re};
  } else if constexpr (KIND == NL_CONV_NONCONS) {
    // -b F[u u_x]  (nonlin_fun/_convection.py:140-171, :204-231)
    cf U[PT], V[PT];
    cf un, vn;
    masked_copy<M, PT>(S, nyq, U, un, j, a.kmax);
    masked_deriv<M, PT>(S, nyq, V, vn, j, a.kmax, a.kscale);
    irfft<M, PT>(U, un, j, lds, tw, twr, a.inv_n);
    irfft<M, PT>(V, vn, j, lds, tw, twr, a.inv_n);
#pragma unroll
    for (int s = 0; s < PT; ++s) U[s] = cf{U[s].re * V[s].re, U[s].im * V[s].im};
    rfft<M, PT>(U, un, j, lds, tw, twr);
    const float c = -a.scale;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const int k = j + s * T;
      const float f = (k <= a.kmax) ? c : 0.f;
      Nn[s] = f * U[s];
    }
    nn = ((M <= a.kmax) ? c : 0.f) * un;
  } else if constexpr (KIND == NL_GRADNORM) {
    // -b/2 F[(u_x)^2 - mean]  (nonlin_fun/_gradient_norm.py:84-101); the mean
    // removal equals zeroing the DC bin of the forward transform.
    cf V[PT];
    cf vn;
    masked_deriv<M, PT>(S, nyq, V, vn, j, a.kmax, a.kscale);
    irfft<M, PT>(V, vn, j, lds, tw, twr, a.inv_n);
#pragma unroll
    for (int s = 0; s < PT; ++s) V[s] = cf{V[s].re * V[s].re, V[s].im * V[s].im};
    rfft<M, PT>(V, vn, j, lds, tw, twr);
    const float c = -a.scale * 0.5f;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const int k = j + s * T;
      float f = (k <= a.kmax) ? c : 0.f;
      if (k == 0 && a.zero_mode_fix) f = 0.f;
      Nn[s] = f * V[s];
    }
    nn = ((M <= a.kmax) ? c : 0.f) * vn;
  } else if constexpr (KIND == NL_POLY) {
    // F[sum_i c_i u^i]  (nonlin_fun/_polynomial.py:64-76), same power loop
    cf U[PT];
    cf un;
    masked_copy<M, PT>(S, nyq, U, un, j, a.kmax);
    irfft<M, PT>(U, un, j, lds, tw, twr, a.inv_n);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      float pr = 1.f, pi = 1.f, ar = 0.f, ai = 0.f;
      for (int i = 0; i < a.npoly; ++i) {
        ar += a.poly[i] * pr;
        ai += a.poly[i] * pi;
        pr *= U[s].re;
        pi *= U[s].im;
      }
      U[s] = cf{ar, ai};
    }
    rfft<M, PT>(U, un, j, lds, tw, twr);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const int k = j + s * T;
      Nn[s] = (k <= a.kmax) ? U[s] : cf{0.f, 0.f};
    }
    nn = (M <= a.kmax) ? un : cf{0.f, 0.f};
  } else if constexpr (KIND == NL_GENERAL) {
    // b0 F[u^2] + (b1/2)(i kappa) F[u^2] + (b2/2) F[u_x^2 - mean]
    // (nonlin_fun/_general_nonlinear.py:89-119 with the sub-functions'
    // signs folded in; F is linear so the u^2 transform is shared)
    cf U[PT], V[PT];
    cf un, vn;
    masked_copy<M, PT>(S, nyq, U, un, j, a.kmax);
    masked_deriv<M, PT>(S, nyq, V, vn, j, a.kmax, a.kscale);
    irfft<M, PT>(U, un, j, lds, tw, twr, a.inv_n);
    irfft<M, PT>(V, vn, j, lds, tw, twr, a.inv_n);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      U[s] = cf{U[s].re * U[s].re, U[s].im * U[s].im};
      V[s] = cf{V[s].re * V[s].re, V[s].im * V[s].im};
    }
    rfft<M, PT>(U, un, j, lds, tw, twr);
    rfft<M, PT>(V, vn, j, lds, tw, twr);
    const float hb1 = 0.5f * a.g1 * a.kscale, hb2 = 0.5f * a.g2;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const int k = j + s * T;
      const bool keep = k <= a.kmax;
      const float ck = hb1 * (float)k;
      const float cg = (k == 0 && a.zero_mode_fix) ? 0.f : hb2;
      cf r = cf{a.g0 * U[s].re - ck * U[s].im + cg * V[s].re, a.g0 * U[s].im + ck * U[s].re + cg * V[s].im};
      Nn[s] = keep ? r : cf{0.f, 0.f};
    }
    {
      const float ck = hb1 * (float)M;
      cf r = cf{a.g0 * un.re - ck * un.im + hb2 * vn.re, a.g0 * un.im + ck * un.re + hb2 * vn.im};
      nn = (M <= a.kmax) ? r : cf{0.f, 0.f};
    }
  }
}

// --------------------------------------------------------------- the kernel
template <int M, int PT, int KIND, int ORDER>
__global__ void __launch_bounds__((M / PT) > 64 ? (M / PT) : 64)
    step1d_kernel(Step1DArgs a) {
  constexpr int T = M / PT;
  constexpr int WG = T > 64 ? T : 64;
  constexpr int G = WG / T;
  constexpr int LDSN = LdsSize<M>::value;
  constexpr int NR = M + 1;
  __shared__ cf lds_all[G * LDSN];

  const int g = threadIdx.x / T;
  const int j = threadIdx.x % T;
  const int t_raw = blockIdx.x * G + g;
  const bool active = t_raw < a.ntrans;
  const int t = active ? t_raw : 0;
  cf* lds = lds_all + g * LDSN;
  const int b = t / a.channels, c = t % a.channels;

  cf S[PT];
  cf nyq = cf{0.f, 0.f};

  // ---- load
  if (a.in_fourier) {
    const cf* in = reinterpret_cast<const cf*>(a.in) + (long long)b * a.in_bstride + (long long)c * NR;
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = in[j + s * T];
    nyq = in[M];
  } else {
    const float2* in = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(a.in) +
                                                       (long long)b * a.in_bstride + (long long)c * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      float2 v = in[j + s * T];
      S[s] = cf{v.x, v.y};
    }
  }

  // operator tables for this transform
  const float* tb = a.tables + ((long long)(a.lin_batch > 1 ? b : 0) * a.lin_channels +
                                (a.lin_channels > 1 ? c : 0)) * (long long)a.ncoef * NR;
  float co[ORDER == 0 ? 1 : (ORDER == 1 ? 2 : (ORDER == 2 ? 3 : (ORDER == 3 ? 7 : 6)))][PT];
  float con[ORDER == 0 ? 1 : (ORDER == 1 ? 2 : (ORDER == 2 ? 3 : (ORDER == 3 ? 7 : 6)))];
  constexpr int NCO = ORDER == 0 ? 1 : (ORDER == 1 ? 2 : (ORDER == 2 ? 3 : (ORDER == 3 ? 7 : 6)));
  if (!a.nonlin_only) {
#pragma unroll
    for (int i = 0; i < NCO; ++i) {
#pragma unroll
      for (int s = 0; s < PT; ++s) co[i][s] = tb[i * NR + j + s * T];
      con[i] = tb[i * NR + M];
    }
  }

  if (!a.in_fourier) rfft<M, PT>(S, nyq, j, lds, a.tw, a.twr);

  if (a.nonlin_only) {
    cf Nn[PT];
    cf nn;
    nl_eval<KIND, M, PT>(S, nyq, Nn, nn, j, lds, a.tw, a.twr, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = Nn[s];
    nyq = nn;
  } else if constexpr (ORDER == 0) {
    // etdrk/_etdrk_0.py:30-34
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = co[0][s] * S[s];
    nyq = con[0] * nyq;
  } else if constexpr (ORDER == 1) {
    // etdrk/_etdrk_1.py:77-82
    cf N0[PT];
    cf n0;
    nl_eval<KIND, M, PT>(S, nyq, N0, n0, j, lds, a.tw, a.twr, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = co[0][s] * S[s] + co[1][s] * N0[s];
    nyq = con[0] * nyq + con[1] * n0;
  } else if constexpr (ORDER == 2) {
    // etdrk/_etdrk_2.py:91-102
    cf N0[PT], N1[PT];
    cf n0, n1;
    nl_eval<KIND, M, PT>(S, nyq, N0, n0, j, lds, a.tw, a.twr, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = co[0][s] * S[s] + co[1][s] * N0[s];
    nyq = con[0] * nyq + con[1] * n0;
    nl_eval<KIND, M, PT>(S, nyq, N1, n1, j, lds, a.tw, a.twr, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = S[s] + co[2][s] * (N1[s] - N0[s]);
    nyq = nyq + con[2] * (n1 - n0);
  } else if constexpr (ORDER == 3) {
    // etdrk/_etdrk_3.py:191-212; tables: E, E/2, c1..c5
    cf N0[PT], Nk[PT], A[PT], U0[PT];
    cf n0, nk, an, u0n;
#pragma unroll
    for (int s = 0; s < PT; ++s) U0[s] = S[s];
    u0n = nyq;
    nl_eval<KIND, M, PT>(S, nyq, N0, n0, j, lds, a.tw, a.twr, a);
    // accumulate the final combination E u + c3 N0 + c4 N1 + c5 N2
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      A[s] = co[0][s] * U0[s] + co[4][s] * N0[s];
      S[s] = co[1][s] * U0[s] + co[2][s] * N0[s];  // stage 1
    }
    an = con[0] * u0n + con[4] * n0;
    nyq = con[1] * u0n + con[2] * n0;
    nl_eval<KIND, M, PT>(S, nyq, Nk, nk, j, lds, a.tw, a.twr, a);  // N1
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      A[s] = A[s] + co[5][s] * Nk[s];
      S[s] = co[0][s] * U0[s] + co[3][s] * (2.f * Nk[s] - N0[s]);  // stage 2
    }
    an = an + con[5] * nk;
    nyq = con[0] * u0n + con[3] * (2.f * nk - n0);
    nl_eval<KIND, M, PT>(S, nyq, Nk, nk, j, lds, a.tw, a.twr, a);  // N2
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = A[s] + co[6][s] * Nk[s];
    nyq = an + con[6] * nk;
  } else {
    // etdrk/_etdrk_4.py:198-224; tables: E, E/2, c1 (= c2 = c3), c4, c5, c6
    cf N0[PT], Nk[PT], A[PT], U0[PT], S1[PT];
    cf n0, nk, an, u0n, s1n;
#pragma unroll
    for (int s = 0; s < PT; ++s) U0[s] = S[s];
    u0n = nyq;
    nl_eval<KIND, M, PT>(S, nyq, N0, n0, j, lds, a.tw, a.twr, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      A[s] = co[0][s] * U0[s] + co[3][s] * N0[s];
      S1[s] = co[1][s] * U0[s] + co[2][s] * N0[s];
    }
    an = con[0] * u0n + con[3] * n0;
    s1n = con[1] * u0n + con[2] * n0;
    nl_eval<KIND, M, PT>(S1, s1n, Nk, nk, j, lds, a.tw, a.twr, a);  // N1
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      A[s] = A[s] + (co[4][s] * 2.f) * Nk[s];
      S[s] = co[1][s] * U0[s] + co[2][s] * Nk[s];  // stage 2
    }
    an = an + (con[4] * 2.f) * nk;
    nyq = con[1] * u0n + con[2] * nk;
    nl_eval<KIND, M, PT>(S, nyq, Nk, nk, j, lds, a.tw, a.twr, a);  // N2
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      A[s] = A[s] + (co[4][s] * 2.f) * Nk[s];
      S[s] = co[1][s] * S1[s] + co[2][s] * (2.f * Nk[s] - N0[s]);  // stage 3
    }
    an = an + (con[4] * 2.f) * nk;
    nyq = con[1] * s1n + con[2] * (2.f * nk - n0);
    nl_eval<KIND, M, PT>(S, nyq, Nk, nk, j, lds, a.tw, a.twr, a);  // N3
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = A[s] + co[5][s] * Nk[s];
    nyq = an + con[5] * nk;
  }

  // ---- store
  if (a.out_fourier) {
    if (active) {
      cf* out = reinterpret_cast<cf*>(a.out) + (long long)b * a.out_bstride + (long long)c * NR;
#pragma unroll
      for (int s = 0; s < PT; ++s) out[j + s * T] = S[s];
      if (j == 0) out[M] = nyq;
    }
  } else {
    irfft<M, PT>(S, nyq, j, lds, a.tw, a.twr, a.inv_n);
    if (active) {
      float2* out = reinterpret_cast<float2*>(reinterpret_cast<float*>(a.out) + (long long)b * a.out_bstride +
                                              (long long)c * 2 * M);
#pragma unroll
      for (int s = 0; s < PT; ++s) out[j + s * T] = make_float2(S[s].re, S[s].im);
    }
  }
}

// Lane-0 Nyquist broadcast is not needed: only lane 0 (k = 0 slot) reads and
// writes `nyq`; other lanes carry a dead copy.

template <int M, int PT, int KIND, int ORDER>
inline hipError_t launch_step1d(const Step1DArgs& a, hipStream_t st) {
  constexpr int T = M / PT;
  constexpr int WG = T > 64 ? T : 64;
  constexpr int G = WG / T;
  const int grid = (a.ntrans + G - 1) / G;
  hipLaunchKernelGGL((step1d_kernel<M, PT, KIND, ORDER>), dim3(grid), dim3(WG), 0, st, a);
  return hipGetLastError();
}

}  // namespace exmi
