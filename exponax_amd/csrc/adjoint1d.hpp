// Adjoint (vector-Jacobian product) of the fused 1-D ETDRK0-2 step.
//
// The reference differentiates the step with jax.grad through lax.scan
// (_utils.py:251). Here the transpose of the step's derivative is written
// out, with the inner products sum(f g) for fields and the Parseval form for
// spectra. Under them:
// - the forward and inverse real transforms are mutual adjoints;
// - a real per-mode factor (E, c1, c2, the dealias mask P) is its own adjoint;
// - a complex E (odd-derivative L) has adjoint conj(E): E is Hermitian in k
//   (E(-k) = conj E(k)), and at the Nyquist mode, whose imaginary part the
//   inverse transform drops, both E and conj(E) act as Re E;
// - i kappa has adjoint -i kappa;
// - the mean removal is self-adjoint.
//
// Nonlinear transposes (mu: cotangent of the nonlinear term; w, w_x: the
// dealiased fields of the linearisation point):
//   convection, conservative   N = -b/2 P i kappa F[w^2]  (_convection.py:104-138, :173-202)
//     J^T mu = -b P F[w F^-1[P (-i kappa) mu]]
//   convection, literal u u_x  N = -b P F[w w_x]          (_convection.py:140-171, :204-231)
//     J^T mu = -b P (F[w_x h] + (-i kappa) F[w h]),  h = F^-1[P mu]
//   gradient norm              N = -b/2 P Z F[w_x^2]       (_gradient_norm.py:84-101)
//     J^T mu = -b P (-i kappa) F[w_x F^-1[P Z mu]]
//   polynomial                 N = P F[p(w)]               (_polynomial.py:64-76)
//     J^T mu = P F[p'(w) F^-1[P mu]]
//   general  N = g0 P F[w^2] + (g1/2) P i kappa F[w^2] + (g2/2) P Z F[w_x^2]
//                                                          (_general_nonlinear.py:89-119)
//     J^T mu = P (F[w (2 g0 h + g1 hd)] + g2 (-i kappa) F[w_x hz]),
//     h = F^-1[P mu], hd = F^-1[P (-i kappa) mu], hz = F^-1[P Z mu]
//
// ETDRK2 (_etdrk_2.py:91-102), cotangent lam of the new state:
//   rho = lam + J(S1)^T (c2 lam),   lam_in = E rho + J(u_hat)^T (c1 rho - c2 lam)
// ETDRK1 (_etdrk_1.py:77-82): lam_in = E lam + J(u_hat)^T (c1 lam); ETDRK0: E lam.
//
// One lane group owns one transform, as in step1d_kernel. The forward stage
// is recomputed from the stored state, and both linearisation points' fields
// stay in registers. Scalings follow step1d.hpp: spectra are carried as 2x and
// physical fields as 2N x, and q2 = 2 / (8 N^2) restores 2x after a product
// (powers of two, exact).
#pragma once
#include "step1d.hpp"

namespace exmi {

struct Vjp1DArgs {
  Step1DArgs s;           // in / in_bstride: the states u (physical float32)
  const float* g;         // cotangent of the new states (B, C, N)
  long long g_bstride;    // floats between trajectories
  float* gout;            // cotangent of u (B, C, N); may alias g
  long long gout_bstride;
};

template <int KIND>
constexpr int lin_fields() {
  return (KIND == NL_CONV_NONCONS || KIND == NL_GENERAL) ? 2 : (KIND == NL_ZERO ? 0 : 1);
}

template <int KIND, int PT>
struct Lin1D {
  static constexpr int NFL = lin_fields<KIND>() > 0 ? lin_fields<KIND>() : 1;
  cf f[NFL][PT];  // physical fields (2N x): CONS w; NONCONS, GENERAL w, w_x; GRADNORM w_x; POLY w
};

// Fields of the linearisation point S (2 u_hat) and, with WANT_N, the
// nonlinear term 2 N_hat in exactly nl_eval's arithmetic.
template <int KIND, bool WANT_N, int M, int PT>
__device__ __forceinline__ void lin_eval(const cf* S, cf nyq, Lin1D<KIND, PT>& L, cf* Nn, cf& nn,
                                         const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& rt,
                                         const Step1DArgs& a) {
  constexpr int T = M / PT;
  constexpr float NN = 2.f * M;
  constexpr float q2 = 2.f / (8.f * NN * NN);
  const int j = ln.j;
  const int kc = a.kmax < M ? a.kmax : M;
  if constexpr (KIND == NL_ZERO) {
    if constexpr (WANT_N) {
#pragma unroll
      for (int s = 0; s < PT; ++s) Nn[s] = cf{0.f, 0.f};
      nn = cf{0.f, 0.f};
    }
  } else if constexpr (KIND == NL_CONV_CONS) {
    cf Z[1][PT];
    cf zn[1];
    masked_copy<M, PT>(S, nyq, Z[0], zn[0], j, a.kmax);
    irfft<M, PT, 1>(Z, zn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) L.f[0][s] = Z[0][s];
    if constexpr (WANT_N) {
#pragma unroll
      for (int s = 0; s < PT; ++s) Z[0][s] = cf{Z[0][s].re * Z[0][s].re, Z[0][s].im * Z[0][s].im};
      rfft<M, PT, 1>(Z, zn, ln, lds, rt, kc);
      const float c = -a.scale * 0.5f * a.kscale * q2;
      const float cj = c * (float)j;
      masked_scale<M, PT>(Z[0], Nn, j, a.kmax, [&](int s, cf x) {
        const float f = fmaf(c, (float)(s * T), cj);
        return cf{-f * x.im, f * x.re};
      });
      const float fn = (M <= a.kmax) ? c * (float)M : 0.f;
      nn = cf{-fn * zn[0].im, fn * zn[0].re};
    }
  } else if constexpr (KIND == NL_CONV_NONCONS) {
    cf U[1][PT], V[1][PT];
    cf un[1], vn[1];
    masked_copy<M, PT>(S, nyq, U[0], un[0], j, a.kmax);
    masked_deriv<M, PT>(S, nyq, V[0], vn[0], j, a.kmax, a.kscale);
    irfft<M, PT, 1>(U, un, ln, lds, rt);
    irfft<M, PT, 1>(V, vn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      L.f[0][s] = U[0][s];
      L.f[1][s] = V[0][s];
    }
    if constexpr (WANT_N) {
#pragma unroll
      for (int s = 0; s < PT; ++s) U[0][s] = cf{U[0][s].re * V[0][s].re, U[0][s].im * V[0][s].im};
      rfft<M, PT, 1>(U, un, ln, lds, rt, kc);
      const float c = -a.scale * q2;
      masked_scale<M, PT>(U[0], Nn, j, a.kmax, [&](int, cf x) { return c * x; });
      nn = ((M <= a.kmax) ? c : 0.f) * un[0];
    }
  } else if constexpr (KIND == NL_POLY) {
    cf Z[1][PT];
    cf zn[1];
    masked_copy<M, PT>(S, nyq, Z[0], zn[0], j, a.kmax);
    irfft<M, PT, 1>(Z, zn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) L.f[0][s] = Z[0][s];
    if constexpr (WANT_N) {
      // nl_eval's power loop on u = (2 N u) / (2 N)
      constexpr float r2n = 1.f / (2.f * NN);
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        const float ur = Z[0][s].re * r2n, ui = Z[0][s].im * r2n;
        float pr = 1.f, pi = 1.f, ar = 0.f, ai = 0.f;
        for (int i = 0; i < a.npoly; ++i) {
          ar += a.poly[i] * pr;
          ai += a.poly[i] * pi;
          pr *= ur;
          pi *= ui;
        }
        Z[0][s] = cf{ar, ai};
      }
      rfft<M, PT, 1>(Z, zn, ln, lds, rt, kc);
      masked_scale<M, PT>(Z[0], Nn, j, a.kmax, [&](int, cf x) { return x; });
      nn = (M <= a.kmax) ? zn[0] : cf{0.f, 0.f};
    }
  } else if constexpr (KIND == NL_GENERAL) {
    cf U[1][PT], V[1][PT];
    cf un[1], vn[1];
    masked_copy<M, PT>(S, nyq, U[0], un[0], j, a.kmax);
    masked_deriv<M, PT>(S, nyq, V[0], vn[0], j, a.kmax, a.kscale);
    irfft<M, PT, 1>(U, un, ln, lds, rt);
    irfft<M, PT, 1>(V, vn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      L.f[0][s] = U[0][s];
      L.f[1][s] = V[0][s];
    }
    if constexpr (WANT_N) {
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        U[0][s] = cf{U[0][s].re * U[0][s].re, U[0][s].im * U[0][s].im};
        V[0][s] = cf{V[0][s].re * V[0][s].re, V[0][s].im * V[0][s].im};
      }
      rfft<M, PT, 1>(U, un, ln, lds, rt, kc);
      rfft<M, PT, 1>(V, vn, ln, lds, rt, kc);
      const float g0 = a.g0 * q2, hb1 = 0.5f * a.g1 * a.kscale * q2, hb2 = 0.5f * a.g2 * q2;
      masked_scale<M, PT>(U[0], Nn, j, a.kmax, [&](int s, cf u) {
        const int k = j + s * T;
        const float ck = hb1 * (float)k;
        const float cg = (k == 0 && a.zero_mode_fix) ? 0.f : hb2;
        const cf v = V[0][s];
        return cf{g0 * u.re - ck * u.im + cg * v.re, g0 * u.im + ck * u.re + cg * v.im};
      });
      const float ckn = hb1 * (float)M;
      nn = (M <= a.kmax) ? cf{g0 * un[0].re - ckn * un[0].im + hb2 * vn[0].re,
                              g0 * un[0].im + ckn * un[0].re + hb2 * vn[0].im}
                         : cf{0.f, 0.f};
    }
  } else {
    static_assert(KIND == NL_GRADNORM, "no 1-D adjoint for this nonlinear kind");
    cf V[1][PT];
    cf vn[1];
    masked_deriv<M, PT>(S, nyq, V[0], vn[0], j, a.kmax, a.kscale);
    irfft<M, PT, 1>(V, vn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) L.f[0][s] = V[0][s];
    if constexpr (WANT_N) {
#pragma unroll
      for (int s = 0; s < PT; ++s) V[0][s] = cf{V[0][s].re * V[0][s].re, V[0][s].im * V[0][s].im};
      rfft<M, PT, 1>(V, vn, ln, lds, rt, kc);
      const float c = -a.scale * 0.5f * q2;
      masked_scale<M, PT>(V[0], Nn, j, a.kmax, [&](int, cf x) { return c * x; });
      if (j == 0 && a.zero_mode_fix) Nn[0] = cf{0.f, 0.f};
      nn = ((M <= a.kmax) ? c : 0.f) * vn[0];
    }
  }
}

// A = 2 J^T mu for X = 2 mu at the linearisation point with fields L.
template <int KIND, int M, int PT>
__device__ __forceinline__ void adj_eval(const Lin1D<KIND, PT>& L, const cf* X, cf xn, cf* A, cf& an,
                                         const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& rt,
                                         const Step1DArgs& a) {
  constexpr int T = M / PT;
  constexpr float NN = 2.f * M;
  constexpr float q2 = 2.f / (8.f * NN * NN);
  const int j = ln.j;
  const int kc = a.kmax < M ? a.kmax : M;
  const float cs = -a.scale * q2;
  const float kn = a.kscale * (float)M;
  if constexpr (KIND == NL_ZERO) {
#pragma unroll
    for (int s = 0; s < PT; ++s) A[s] = cf{0.f, 0.f};
    an = cf{0.f, 0.f};
  } else if constexpr (KIND == NL_CONV_CONS) {
    // h = F^-1[P (-i kappa) mu]; A = cs P F[w h]
    cf Z[1][PT];
    cf zn[1];
    masked_deriv<M, PT>(X, xn, Z[0], zn[0], j, a.kmax, a.kscale);
#pragma unroll
    for (int s = 0; s < PT; ++s) Z[0][s] = cf{-Z[0][s].re, -Z[0][s].im};
    zn[0] = cf{-zn[0].re, -zn[0].im};
    irfft<M, PT, 1>(Z, zn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) Z[0][s] = cf{L.f[0][s].re * Z[0][s].re, L.f[0][s].im * Z[0][s].im};
    rfft<M, PT, 1>(Z, zn, ln, lds, rt, kc);
    masked_scale<M, PT>(Z[0], A, j, a.kmax, [&](int, cf x) { return cs * x; });
    an = ((M <= a.kmax) ? cs : 0.f) * zn[0];
  } else if constexpr (KIND == NL_CONV_NONCONS) {
    // h = F^-1[P mu]; A = cs P (F[w_x h] + (-i kappa) F[w h])
    cf Z[2][PT];
    cf zn[2];
    masked_copy<M, PT>(X, xn, Z[0], zn[0], j, a.kmax);
    {
      cf H[1][PT];
      cf hn[1] = {zn[0]};
#pragma unroll
      for (int s = 0; s < PT; ++s) H[0][s] = Z[0][s];
      irfft<M, PT, 1>(H, hn, ln, lds, rt);
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        Z[0][s] = cf{L.f[1][s].re * H[0][s].re, L.f[1][s].im * H[0][s].im};
        Z[1][s] = cf{L.f[0][s].re * H[0][s].re, L.f[0][s].im * H[0][s].im};
      }
    }
    rfft<M, PT, 1>(Z, zn, ln, lds, rt, kc);
    rfft<M, PT, 1>(Z + 1, zn + 1, ln, lds, rt, kc);
    const float cj = cs * a.kscale * (float)j;
    const float ck = cs * a.kscale;
    masked_scale<M, PT>(Z[0], A, j, a.kmax, [&](int s, cf x) {
      const float f = fmaf(ck, (float)(s * T), cj);  // cs * kappa_k
      const cf y = Z[1][s];
      return cf{fmaf(cs, x.re, f * y.im), fmaf(cs, x.im, -f * y.re)};
    });
    an = (M <= a.kmax) ? cf{fmaf(cs, zn[0].re, cs * kn * zn[1].im), fmaf(cs, zn[0].im, -cs * kn * zn[1].re)}
                       : cf{0.f, 0.f};
  } else if constexpr (KIND == NL_POLY) {
    // h = F^-1[P mu] (2N h from X = 2 mu); A = P F[p'(w) h]: the forward
    // transform of p'(w) (2N h) gives 4N F[p' h], so A = 2 J^T mu needs 1/(2N)
    constexpr float r2n = 1.f / (2.f * NN);
    cf Z[1][PT];
    cf zn[1];
    masked_copy<M, PT>(X, xn, Z[0], zn[0], j, a.kmax);
    irfft<M, PT, 1>(Z, zn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const float wr = L.f[0][s].re * r2n, wi = L.f[0][s].im * r2n;
      float pr = 1.f, pi = 1.f, dr = 0.f, di = 0.f;
      for (int i = 1; i < a.npoly; ++i) {
        dr += (float)i * a.poly[i] * pr;
        di += (float)i * a.poly[i] * pi;
        pr *= wr;
        pi *= wi;
      }
      Z[0][s] = cf{dr * Z[0][s].re, di * Z[0][s].im};
    }
    rfft<M, PT, 1>(Z, zn, ln, lds, rt, kc);
    masked_scale<M, PT>(Z[0], A, j, a.kmax, [&](int, cf x) { return r2n * x; });
    an = ((M <= a.kmax) ? r2n : 0.f) * zn[0];
    (void)cs;
    (void)kn;
  } else if constexpr (KIND == NL_GENERAL) {
    cf H[1][PT], D[1][PT], Q[1][PT];
    cf hn[1], dn[1], qn[1];
    masked_copy<M, PT>(X, xn, H[0], hn[0], j, a.kmax);
    masked_deriv<M, PT>(X, xn, D[0], dn[0], j, a.kmax, a.kscale);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      D[0][s] = cf{-D[0][s].re, -D[0][s].im};
      Q[0][s] = H[0][s];
    }
    dn[0] = cf{-dn[0].re, -dn[0].im};
    qn[0] = hn[0];
    if (j == 0 && a.zero_mode_fix) Q[0][0] = cf{0.f, 0.f};
    irfft<M, PT, 1>(H, hn, ln, lds, rt);
    irfft<M, PT, 1>(D, dn, ln, lds, rt);
    irfft<M, PT, 1>(Q, qn, ln, lds, rt);
    const float c0 = 2.f * a.g0 * q2, c1 = a.g1 * q2, c2 = a.g2 * q2;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf w = L.f[0][s], wx = L.f[1][s];
      H[0][s] = cf{w.re * fmaf(c0, H[0][s].re, c1 * D[0][s].re), w.im * fmaf(c0, H[0][s].im, c1 * D[0][s].im)};
      Q[0][s] = cf{wx.re * Q[0][s].re, wx.im * Q[0][s].im};
    }
    rfft<M, PT, 1>(H, hn, ln, lds, rt, kc);
    rfft<M, PT, 1>(Q, qn, ln, lds, rt, kc);
    const float cj = c2 * a.kscale * (float)j, ck = c2 * a.kscale;
    masked_scale<M, PT>(H[0], A, j, a.kmax, [&](int s, cf x) {
      const float f = fmaf(ck, (float)(s * T), cj);  // g2 q2 kappa_k
      const cf y = Q[0][s];
      return cf{fmaf(f, y.im, x.re), fmaf(-f, y.re, x.im)};  // x + f (-i) y
    });
    const float fn = c2 * kn;
    an = (M <= a.kmax) ? cf{fmaf(fn, qn[0].im, hn[0].re), fmaf(-fn, qn[0].re, hn[0].im)} : cf{0.f, 0.f};
    (void)cs;
  } else {
    static_assert(KIND == NL_GRADNORM, "no 1-D adjoint for this nonlinear kind");
    // h = F^-1[P Z mu]; A = cs P (-i kappa) F[w_x h]
    cf Z[1][PT];
    cf zn[1];
    masked_copy<M, PT>(X, xn, Z[0], zn[0], j, a.kmax);
    if (j == 0 && a.zero_mode_fix) Z[0][0] = cf{0.f, 0.f};
    irfft<M, PT, 1>(Z, zn, ln, lds, rt);
#pragma unroll
    for (int s = 0; s < PT; ++s) Z[0][s] = cf{L.f[0][s].re * Z[0][s].re, L.f[0][s].im * Z[0][s].im};
    rfft<M, PT, 1>(Z, zn, ln, lds, rt, kc);
    const float cj = cs * a.kscale * (float)j;
    const float ck = cs * a.kscale;
    masked_scale<M, PT>(Z[0], A, j, a.kmax, [&](int s, cf x) {
      const float f = fmaf(ck, (float)(s * T), cj);
      return cf{f * x.im, -f * x.re};
    });
    const float fn = (M <= a.kmax) ? cs * kn : 0.f;
    an = cf{fn * zn[0].im, -fn * zn[0].re};
  }
}

template <int M, int PT, int KIND, int ORDER, bool CPLX = false>
__global__ void __launch_bounds__((M / PT) > 64 ? (M / PT) : 64) vjp1d_kernel(Vjp1DArgs va) {
  const Step1DArgs& a = va.s;
  constexpr int T = M / PT;
  constexpr int G = Slots<M, PT, 1>::G;
  constexpr int LDSN = LdsSize<M>::value;
  constexpr int NR = M + 1;
  constexpr int NRE = ncoef_of(ORDER);
  constexpr int NCO = NRE + (CPLX ? 1 : 0);  // Im E follows the real coefficients
  using CL = CoefLds<M, PT, NCO>;
  constexpr float NN = 2.f * M;
  constexpr float r2n = 1.f / (2.f * NN);
  __shared__ cf lds_all[G * LDSN];
  __shared__ float lds_coef[G * CL::FLOATS];

  const int g = threadIdx.x / T;
  const int j = threadIdx.x % T;
  const int t_raw = blockIdx.x * G + g;
  const bool active = t_raw < a.ntrans;
  const int t = active ? t_raw : 0;
  cf* lds = lds_all + g * LDSN;
  const CL co{lds_coef + g * CL::FLOATS, j};
  // E x (the forward stage) and conj(E) x (the transpose)
  auto emul = [&](int s, cf x) -> cf {
    const float er = co(0, 0, s);
    if constexpr (CPLX) {
      const float ei = co(0, NRE, s);
      return cf{fmaf(er, x.re, -ei * x.im), fmaf(er, x.im, ei * x.re)};
    } else {
      return er * x;
    }
  };
  auto etmul = [&](int s, cf x) -> cf {
    const float er = co(0, 0, s);
    if constexpr (CPLX) {
      const float ei = co(0, NRE, s);
      return cf{fmaf(er, x.re, ei * x.im), fmaf(er, x.im, -ei * x.re)};
    } else {
      return er * x;
    }
  };
  const int b = t / a.channels, c = t % a.channels;
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> rt;
  rt.load(j, RealTables{a.tw, a.fw});

  // state and cotangent (physical float32)
  cf U[1][PT], Lm[1][PT];
  cf un[1] = {cf{0.f, 0.f}}, lmn[1] = {cf{0.f, 0.f}};
  {
    const float2* in = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(a.in) +
                                                       (long long)b * a.in_bstride + (long long)c * 2 * M);
    const float2* gi = reinterpret_cast<const float2*>(va.g + (long long)b * va.g_bstride + (long long)c * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const float2 v = in[j + s * T];
      U[0][s] = cf{v.x, v.y};
      const float2 w = gi[j + s * T];
      Lm[0][s] = cf{w.x, w.y};
    }
  }
  {
    const float* tb = a.tables + ((long long)(a.lin_batch > 1 ? b : 0) * a.lin_channels +
                                  (a.lin_channels > 1 ? c : 0)) * (long long)NCO * NR;
    CoefRegs<NCO, PT, 1> cr;
    cr.template load<M, NATURAL>(0, tb, j);
    rfft<M, PT, 1>(U, un, ln, lds, rt);
#pragma unroll
    for (int i = 0; i < NCO; ++i)
#pragma unroll
      for (int s = 0; s <= PT; ++s) co.put(0, i, s, cr(0, i, s));
  }
  rfft<M, PT, 1>(Lm, lmn, ln, lds, rt);  // 2 lam

  cf O[1][PT];
  cf on[1];
  if constexpr (ORDER == 0) {
#pragma unroll
    for (int s = 0; s < PT; ++s) O[0][s] = etmul(s, Lm[0][s]);
    on[0] = etmul(PT, lmn[0]);
  } else if constexpr (ORDER == 1) {
    Lin1D<KIND, PT> L0;
    cf dum[PT], dn;
    lin_eval<KIND, false, M, PT>(U[0], un[0], L0, dum, dn, ln, lds, rt, a);
    cf X[PT], A[PT], an;
#pragma unroll
    for (int s = 0; s < PT; ++s) X[s] = co(0, 1, s) * Lm[0][s];
    const cf xn = co(0, 1, PT) * lmn[0];
    adj_eval<KIND, M, PT>(L0, X, xn, A, an, ln, lds, rt, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) O[0][s] = etmul(s, Lm[0][s]) + A[s];
    on[0] = etmul(PT, lmn[0]) + an;
  } else {
    static_assert(ORDER == 2, "1-D adjoint implements ETDRK orders 0-2");
    Lin1D<KIND, PT> L0, L1;
    cf N0[PT], n0;
    lin_eval<KIND, true, M, PT>(U[0], un[0], L0, N0, n0, ln, lds, rt, a);
    cf S1[PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) S1[s] = emul(s, U[0][s]) + co(0, 1, s) * N0[s];
    const cf s1n = emul(PT, un[0]) + co(0, 1, PT) * n0;
    lin_eval<KIND, false, M, PT>(S1, s1n, L1, N0, n0, ln, lds, rt, a);
    // rho = lam + J(S1)^T (c2 lam)
    cf X[PT], A[PT], an;
#pragma unroll
    for (int s = 0; s < PT; ++s) X[s] = co(0, 2, s) * Lm[0][s];
    cf xn = co(0, 2, PT) * lmn[0];
    adj_eval<KIND, M, PT>(L1, X, xn, A, an, ln, lds, rt, a);
    cf R[PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      R[s] = Lm[0][s] + A[s];
      X[s] = co(0, 1, s) * R[s] - co(0, 2, s) * Lm[0][s];
    }
    const cf rn = lmn[0] + an;
    xn = co(0, 1, PT) * rn - co(0, 2, PT) * lmn[0];
    adj_eval<KIND, M, PT>(L0, X, xn, A, an, ln, lds, rt, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) O[0][s] = etmul(s, R[s]) + A[s];
    on[0] = etmul(PT, rn) + an;
  }
  irfft<M, PT, 1>(O, on, ln, lds, rt);
  if (active) {
    float2* out = reinterpret_cast<float2*>(va.gout + (long long)b * va.gout_bstride + (long long)c * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) out[j + s * T] = make_float2(O[0][s].re * r2n, O[0][s].im * r2n);
  }
}

template <int M, int PT, int KIND, int ORDER, bool CPLX = false>
inline hipError_t launch_vjp1d(const Vjp1DArgs& va, hipStream_t st) {
  constexpr int WG = Slots<M, PT, 1>::WG;
  constexpr int G = Slots<M, PT, 1>::G;
  const int grid = (va.s.ntrans + G - 1) / G;
  hipLaunchKernelGGL((vjp1d_kernel<M, PT, KIND, ORDER, CPLX>), dim3(grid), dim3(WG), 0, st, va);
  return hipGetLastError();
}

}  // namespace exmi
