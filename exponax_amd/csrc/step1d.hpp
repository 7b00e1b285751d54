// Fused 1-D ETDRK0-4 pseudo-spectral step for gfx950.
//
// One launch advances every (trajectory, channel) transform of a batch by one
// step (or evaluates the nonlinear term only). A transform is owned by T =
// M/PT lanes and stays in registers for the whole step: the physical state is
// read once (float2 pairs, coalesced), transformed, every nonlinear
// evaluation (dealias, derivative multiply, products, forward transform, post
// dealias) and ETDRK stage combination happens on registers with LDS used
// for the FFT exchanges and as lane-private storage of the operator tables,
// and the new state is written once.
//
// Scaling: the forward transform yields 2 X and the inverse yields N x (see
// fft1d.hpp). Inside the kernel the spectral state is carried as 2 u_hat and
// every nonlinear term as 2 N_hat, so the ETDRK combinations are unchanged;
// the nonlinear coefficients absorb the remaining powers of two (exact).
//
// Semantics follow exponax 0.2.0: BaseStepper.step (_base_stepper.py:201-220),
// ETDRK0-4 step_fourier (etdrk/_etdrk_0.py:30-34, _etdrk_1.py:77-82,
// _etdrk_2.py:91-102, _etdrk_3.py:191-212, _etdrk_4.py:198-224) and the
// nonlinear functions of nonlin_fun/ (see nl_eval below).
#pragma once
#include <climits>
#include <type_traits>

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
  const cf* fw;          // -i exp(-2 pi i k / N), k < M
  float kscale;          // 2 pi / L (float32, as the reference's derivative operator)
  int kmax;              // dealias: keep |k| <= kmax (kmax >= M: keep all)
  float scale;           // b of convection / gradient norm
  float g0, g1, g2;      // general nonlinear scales (b0, b1, b2)
  int zero_mode_fix;
  int npoly;
  float poly[EXMI_MAX_POLY];
  int poly_lap;          // 1: the polynomial's transform times poly_lap_c k^2 (Cahn-Hilliard)
  float poly_lap_c;      // -scale (2 pi / L)^2: scale * Laplace in units of k^2
  float feed, kill;      // Gray-Scott rates
};


constexpr int ncoef_of(int order) {
  return order == 0 ? 1 : (order == 1 ? 2 : (order == 2 ? 3 : (order == 3 ? 7 : 6)));
}
// table entries per mode, with Im E (and Im E_half) for a complex linear operator
constexpr int ncoef_of(int order, bool cplx) { return ncoef_of(order) + (cplx ? (order >= 3 ? 2 : 1) : 0); }

// ----------------------------------------------------------------- helpers
// Dealias mask classes of slot s for a uniform kmax: every mode kept, every
// mode dropped, or a per-lane boundary. The branches are uniform across the
// wave, so fully kept or dropped slots cost no VALU selects. Slot s of lane j
// holds mode Modes<M, PT, LAY>::k(j, s) (fft1d.hpp); the extra register holds
// mode Modes::kextra.
template <int M, int PT, int LAY>
__device__ __forceinline__ bool slot_kept(int s, int kmax) { return Modes<M, PT, LAY>::khi(s) <= kmax; }
template <int M, int PT, int LAY>
__device__ __forceinline__ bool slot_dropped(int s, int kmax) { return Modes<M, PT, LAY>::klo(s) > kmax; }

// c * k for slot s of lane j: cj = c * j, cjm = c * (M - j) (exact integer
// mode numbers; one fused multiply-add per slot)
template <int M, int PT, int LAY>
__device__ __forceinline__ float scaled_mode(float c, float cj, float cjm, int s) {
  constexpr int T = M / PT, H = PT / 2;
  if constexpr (LAY == NATURAL) return fmaf(c, (float)(s * T), cj);
  else return s < H ? fmaf(c, (float)(s * T), cj) : fmaf(-c, (float)((s - H) * T), cjm);
}

// Z = mask * S (dealias before an inverse transform)
template <int M, int PT, int LAY = NATURAL>
__device__ __forceinline__ void masked_copy(const cf* S, cf nyq, cf* Z, cf& zn, int j, int kmax) {
  using MD = Modes<M, PT, LAY>;
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    if (slot_kept<M, PT, LAY>(s, kmax)) {
      Z[s] = S[s];
    } else if (slot_dropped<M, PT, LAY>(s, kmax)) {
      Z[s] = cf{0.f, 0.f};
    } else {
      Z[s] = (MD::k(j, s) <= kmax) ? S[s] : cf{0.f, 0.f};
    }
  }
  zn = (MD::kextra <= kmax) ? nyq : cf{0.f, 0.f};
}

// Z = mask * (i kappa) * S
template <int M, int PT, int LAY = NATURAL>
__device__ __forceinline__ void masked_deriv(const cf* S, cf nyq, cf* Z, cf& zn, int j, int kmax,
                                             float kscale) {
  using MD = Modes<M, PT, LAY>;
  const float kj = kscale * (float)j, kjm = kscale * (float)(M - j);
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    if (slot_dropped<M, PT, LAY>(s, kmax)) {
      Z[s] = cf{0.f, 0.f};
      continue;
    }
    const float kap = scaled_mode<M, PT, LAY>(kscale, kj, kjm, s);  // kscale * k
    const cf z = cf{-kap * S[s].im, kap * S[s].re};
    if (slot_kept<M, PT, LAY>(s, kmax)) {
      Z[s] = z;
    } else {
      Z[s] = (MD::k(j, s) <= kmax) ? z : cf{0.f, 0.f};
    }
  }
  const float kn = kscale * (float)MD::kextra;
  zn = (MD::kextra <= kmax) ? cf{-kn * nyq.im, kn * nyq.re} : cf{0.f, 0.f};
}

// Nn = mask * f * X for a real per-mode factor f(k) (dealias after a forward
// transform). Fully dropped slots are already zero from rfft's kcut.
template <int M, int PT, int LAY = NATURAL, class F>
__device__ __forceinline__ void masked_scale(const cf* X, cf* Nn, int j, int kmax, F f) {
  using MD = Modes<M, PT, LAY>;
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    if (slot_dropped<M, PT, LAY>(s, kmax)) {
      Nn[s] = cf{0.f, 0.f};
      continue;
    }
    const cf y = f(s, X[s]);
    if (slot_kept<M, PT, LAY>(s, kmax)) {
      Nn[s] = y;
    } else {
      Nn[s] = (MD::k(j, s) <= kmax) ? y : cf{0.f, 0.f};
    }
  }
}

// ------------------------------------------------------ nonlinear functions
// Input: 2 u_hat (lane-strided + Nyquist) for NT transforms. Output: 2 N_hat,
// dealiased. With the transform scalings, a product of two inverse
// transforms of 2 u_hat comes back from the forward transform as
// 8 N^2 F[product]; the factor q2 = 2 / (8 N^2) restores 2 N_hat (powers of
// two, exact).
template <int KIND, int M, int PT, int NT, int LAY = NATURAL>
__device__ __forceinline__ void nl_eval(const cf (*S)[PT], const cf* nyq, cf (*Nn)[PT], cf* nn,
                                        const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& rt,
                                        const Step1DArgs& a) {
  using MD = Modes<M, PT, LAY>;
  constexpr int KX = MD::kextra;
  constexpr float NN = 2.f * M;
  constexpr float q2 = 2.f / (8.f * NN * NN);
  const int j = ln.j;
  const int kc = a.kmax < M ? a.kmax : M;  // forward transforms are dealiased
  if constexpr (KIND == NL_ZERO) {
    // nonlin_fun/_zero.py:33-38
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
      for (int s = 0; s < PT; ++s) Nn[n][s] = cf{0.f, 0.f};
      nn[n] = cf{0.f, 0.f};
    }
  } else if constexpr (KIND == NL_CONV_CONS) {
    // -b/2 (i kappa) F[u^2]  (nonlin_fun/_convection.py:104-138, :173-202)
    cf Z[NT][PT];
    cf zn[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) masked_copy<M, PT, LAY>(S[n], nyq[n], Z[n], zn[n], j, a.kmax);
    irfft_l<LAY, M, PT, NT>(Z, zn, ln, lds, rt, a.kmax);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) Z[n][s] = cf{Z[n][s].re * Z[n][s].re, Z[n][s].im * Z[n][s].im};
    rfft_l<LAY, M, PT, NT>(Z, zn, ln, lds, rt, kc);
    const float c = -a.scale * 0.5f * a.kscale * q2;
    const float cj = c * (float)j, cjm = c * (float)(M - j);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_scale<M, PT, LAY>(Z[n], Nn[n], j, a.kmax, [&](int s, cf x) {
        const float f = scaled_mode<M, PT, LAY>(c, cj, cjm, s);  // c * k
        return cf{-f * x.im, f * x.re};
      });
      const float fn = (KX <= a.kmax) ? c * (float)KX : 0.f;
      nn[n] = cf{-fn * zn[n].im, fn * zn[n].re};
    }
  } else if constexpr (KIND == NL_CONV_NONCONS) {
    // -b F[u u_x]  (nonlin_fun/_convection.py:140-171, :204-231)
    cf U[NT][PT], V[NT][PT];
    cf un[NT], vn[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_copy<M, PT, LAY>(S[n], nyq[n], U[n], un[n], j, a.kmax);
      masked_deriv<M, PT, LAY>(S[n], nyq[n], V[n], vn[n], j, a.kmax, a.kscale);
    }
    irfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, a.kmax);
    irfft_l<LAY, M, PT, NT>(V, vn, ln, lds, rt, a.kmax);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) U[n][s] = cf{U[n][s].re * V[n][s].re, U[n][s].im * V[n][s].im};
    rfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, kc);
    const float c = -a.scale * q2;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_scale<M, PT, LAY>(U[n], Nn[n], j, a.kmax, [&](int, cf x) { return c * x; });
      nn[n] = ((KX <= a.kmax) ? c : 0.f) * un[n];
    }
  } else if constexpr (KIND == NL_GRADNORM) {
    // -b/2 F[(u_x)^2 - mean]  (nonlin_fun/_gradient_norm.py:84-101); the mean
    // removal equals zeroing the DC bin of the forward transform.
    cf V[NT][PT];
    cf vn[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) masked_deriv<M, PT, LAY>(S[n], nyq[n], V[n], vn[n], j, a.kmax, a.kscale);
    irfft_l<LAY, M, PT, NT>(V, vn, ln, lds, rt, a.kmax);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) V[n][s] = cf{V[n][s].re * V[n][s].re, V[n][s].im * V[n][s].im};
    rfft_l<LAY, M, PT, NT>(V, vn, ln, lds, rt, kc);
    const float c = -a.scale * 0.5f * q2;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_scale<M, PT, LAY>(V[n], Nn[n], j, a.kmax, [&](int, cf x) { return c * x; });
      if (j == 0 && a.zero_mode_fix) Nn[n][0] = cf{0.f, 0.f};
      nn[n] = ((KX <= a.kmax) ? c : 0.f) * vn[n];
    }
  } else if constexpr (KIND == NL_POLY) {
    // F[sum_i c_i u^i]  (nonlin_fun/_polynomial.py:64-76), same power loop;
    // the inverse transform of 2 u_hat is 2 N u, rescaled first.
    cf U[NT][PT];
    cf un[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) masked_copy<M, PT, LAY>(S[n], nyq[n], U[n], un[n], j, a.kmax);
    irfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, a.kmax);
    constexpr float r2n = 1.f / (2.f * NN);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        const float ur = U[n][s].re * r2n, ui = U[n][s].im * r2n;
        float pr = 1.f, pi = 1.f, ar = 0.f, ai = 0.f;
        for (int i = 0; i < a.npoly; ++i) {
          ar += a.poly[i] * pr;
          ai += a.poly[i] * pi;
          pr *= ur;
          pi *= ui;
        }
        U[n][s] = cf{ar, ai};
      }
    rfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, kc);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      if (a.poly_lap) {
        masked_scale<M, PT, LAY>(U[n], Nn[n], j, a.kmax, [&](int s, cf x) {
          const float k = (float)MD::k(j, s);
          return (a.poly_lap_c * k * k) * x;
        });
        nn[n] = (KX <= a.kmax) ? (a.poly_lap_c * (float)KX * (float)KX) * un[n] : cf{0.f, 0.f};
      } else {
        masked_scale<M, PT, LAY>(U[n], Nn[n], j, a.kmax, [&](int, cf x) { return x; });
        nn[n] = (KX <= a.kmax) ? un[n] : cf{0.f, 0.f};
      }
    }
  } else if constexpr (KIND == NL_GRAY_SCOTT || KIND == NL_BZ) {
    // the NT transforms of a group are the C channels of one trajectory
    static_assert(NT == (KIND == NL_GRAY_SCOTT ? 2 : 3), "one trajectory's channels per lane group");
    cf U[NT][PT];
    cf un[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) masked_copy<M, PT, LAY>(S[n], nyq[n], U[n], un[n], j, a.kmax);
    irfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, a.kmax);
    constexpr float r2n = 1.f / (2.f * NN);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      // real and imaginary parts hold two grid points each
      float v[NT][2];
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        v[n][0] = U[n][s].re * r2n;
        v[n][1] = U[n][s].im * r2n;
      }
      float r[NT][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if constexpr (KIND == NL_GRAY_SCOTT) {
          gray_scott(v[0][h], v[1][h], a.feed, a.kill, r[0][h], r[1][h]);
        } else {
          belousov_zhabotinsky(v[0][h], v[1][h], v[2][h], r[0][h], r[1][h], r[2][h]);
        }
      }
#pragma unroll
      for (int n = 0; n < NT; ++n) U[n][s] = cf{r[n][0], r[n][1]};
    }
    rfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, kc);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_scale<M, PT, LAY>(U[n], Nn[n], j, a.kmax, [&](int, cf x) { return x; });
      nn[n] = (KX <= a.kmax) ? un[n] : cf{0.f, 0.f};
    }
  } else if constexpr (KIND == NL_GENERAL) {
    // b0 F[u^2] + (b1/2)(i kappa) F[u^2] + (b2/2) F[u_x^2 - mean]
    // (nonlin_fun/_general_nonlinear.py:89-119 with the sub-functions'
    // signs folded in; F is linear so the u^2 transform is shared)
    cf U[NT][PT], V[NT][PT];
    cf un[NT], vn[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_copy<M, PT, LAY>(S[n], nyq[n], U[n], un[n], j, a.kmax);
      masked_deriv<M, PT, LAY>(S[n], nyq[n], V[n], vn[n], j, a.kmax, a.kscale);
    }
    irfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, a.kmax);
    irfft_l<LAY, M, PT, NT>(V, vn, ln, lds, rt, a.kmax);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        U[n][s] = cf{U[n][s].re * U[n][s].re, U[n][s].im * U[n][s].im};
        V[n][s] = cf{V[n][s].re * V[n][s].re, V[n][s].im * V[n][s].im};
      }
    rfft_l<LAY, M, PT, NT>(U, un, ln, lds, rt, kc);
    rfft_l<LAY, M, PT, NT>(V, vn, ln, lds, rt, kc);
    const float g0 = a.g0 * q2, hb1 = 0.5f * a.g1 * a.kscale * q2, hb2 = 0.5f * a.g2 * q2;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        const int k = MD::k(j, s);
        const bool keep = k <= a.kmax;
        const float ck = hb1 * (float)k;
        const float cg = (k == 0 && a.zero_mode_fix) ? 0.f : hb2;
        const cf u = U[n][s], v = V[n][s];
        cf r = cf{g0 * u.re - ck * u.im + cg * v.re, g0 * u.im + ck * u.re + cg * v.im};
        Nn[n][s] = keep ? r : cf{0.f, 0.f};
      }
      const float ck = hb1 * (float)KX;
      const cf u = un[n], v = vn[n];
      cf r = cf{g0 * u.re - ck * u.im + hb2 * v.re, g0 * u.im + ck * u.re + hb2 * v.im};
      nn[n] = (KX <= a.kmax) ? r : cf{0.f, 0.f};
    }
  }
}

// ------------------------------------------------------- operator tables
// Accessors co(n, i, s): coefficient i of transform n at slot s (s = PT is
// the Nyquist entry, meaningful in lane 0).
template <int NCO, int PT, int NT>
struct CoefRegs {
  float v[NT][NCO][PT + 1];
  __device__ __forceinline__ float operator()(int n, int i, int s) const { return v[n][i][s]; }
  template <int M, int LAY>
  __device__ __forceinline__ void load(int n, const float* __restrict__ tb, int j) {
    using MD = Modes<M, PT, LAY>;
#pragma unroll
    for (int i = 0; i < NCO; ++i) {
#pragma unroll
      for (int s = 0; s < PT; ++s) v[n][i][s] = tb[i * (M + 1) + MD::k(j, s)];
      v[n][i][PT] = tb[i * (M + 1) + MD::kextra];
    }
  }
};

// Lane-private LDS copy (frees the registers between the table load and its
// uses).
template <int M, int PT, int NCO>
struct CoefLds {
  static constexpr int T = M / PT;
  static constexpr int FLOATS = NCO * (PT + 1) * T;  // per transform
  float* base;
  int j;
  __device__ __forceinline__ float operator()(int n, int i, int s) const {
    return base[((n * NCO + i) * (PT + 1) + s) * T + j];
  }
  __device__ __forceinline__ void put(int n, int i, int s, float x) const {
    base[((n * NCO + i) * (PT + 1) + s) * T + j] = x;
  }
};

// ------------------------------------------------------- one ETDRK step
// S: 2 u_hat of NT transforms (in/out). etdrk/_etdrk_0.py:30-34,
// _etdrk_1.py:77-82, _etdrk_2.py:91-102, _etdrk_3.py:191-212,
// _etdrk_4.py:198-224. Table order: ETDRK0 [E]; ETDRK1 [E, c1];
// ETDRK2 [E, c1, c2]; ETDRK3 [E, E/2, c1..c5]; ETDRK4 [E, E/2, c1, c4, c5, c6].
template <int KIND, int ORDER, int M, int PT, int NT, class Coef, int LAY = NATURAL, bool CPLX = false>
__device__ __forceinline__ void etdrk_step(cf (*S)[PT], cf* nyq, const Coef& co, const Lane<M, PT>& ln, cf* lds,
                                           const TwRegs<M, PT>& rt, const Step1DArgs& a) {
  // E x and E_half x: real, or complex with Im E (and Im E_half for orders
  // 3-4) stored after the real coefficients
  constexpr int NRE = ncoef_of(ORDER);
  auto cmul = [&](int n, int i, int s, cf x) -> cf {
    const float er = co(n, i, s);
    if constexpr (CPLX) {
      const float ei = co(n, NRE + i, s);
      return cf{fmaf(er, x.re, -ei * x.im), fmaf(er, x.im, ei * x.re)};
    } else {
      return er * x;
    }
  };
  auto emul = [&](int n, int s, cf x) -> cf { return cmul(n, 0, s, x); };
  auto hmul = [&](int n, int s, cf x) -> cf { return cmul(n, 1, s, x); };
#define EXMI_EACH(body)                   \
  _Pragma("unroll") for (int n = 0; n < NT; ++n) { \
    _Pragma("unroll") for (int s = 0; s < PT; ++s) { body; } }
  if constexpr (ORDER == 0) {
    EXMI_EACH(S[n][s] = emul(n, s, S[n][s]));
#pragma unroll
    for (int n = 0; n < NT; ++n) nyq[n] = emul(n, PT, nyq[n]);
  } else if constexpr (ORDER == 1) {
    cf N0[NT][PT];
    cf n0[NT];
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, N0, n0, ln, lds, rt, a);
    EXMI_EACH(S[n][s] = emul(n, s, S[n][s]) + co(n, 1, s) * N0[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) nyq[n] = emul(n, PT, nyq[n]) + co(n, 1, PT) * n0[n];
  } else if constexpr (ORDER == 2) {
    cf N0[NT][PT], N1[NT][PT];
    cf n0[NT], n1[NT];
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, N0, n0, ln, lds, rt, a);
    EXMI_EACH(S[n][s] = emul(n, s, S[n][s]) + co(n, 1, s) * N0[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) nyq[n] = emul(n, PT, nyq[n]) + co(n, 1, PT) * n0[n];
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, N1, n1, ln, lds, rt, a);
    EXMI_EACH(S[n][s] = S[n][s] + co(n, 2, s) * (N1[n][s] - N0[n][s]));
#pragma unroll
    for (int n = 0; n < NT; ++n) nyq[n] = nyq[n] + co(n, 2, PT) * (n1[n] - n0[n]);
  } else if constexpr (ORDER == 3) {
    cf N0[NT][PT], Nk[NT][PT], A[NT][PT], U0[NT][PT];
    cf n0[NT], nk[NT], an[NT], u0n[NT];
    EXMI_EACH(U0[n][s] = S[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) u0n[n] = nyq[n];
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, N0, n0, ln, lds, rt, a);
    // accumulate the final combination E u + c3 N0 + c4 N1 + c5 N2
    EXMI_EACH(A[n][s] = emul(n, s, U0[n][s]) + co(n, 4, s) * N0[n][s];
              S[n][s] = hmul(n, s, U0[n][s]) + co(n, 2, s) * N0[n][s]);  // stage 1
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      an[n] = emul(n, PT, u0n[n]) + co(n, 4, PT) * n0[n];
      nyq[n] = hmul(n, PT, u0n[n]) + co(n, 2, PT) * n0[n];
    }
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, Nk, nk, ln, lds, rt, a);  // N1
    EXMI_EACH(A[n][s] = A[n][s] + co(n, 5, s) * Nk[n][s];
              S[n][s] = emul(n, s, U0[n][s]) + co(n, 3, s) * (2.f * Nk[n][s] - N0[n][s]));  // stage 2
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      an[n] = an[n] + co(n, 5, PT) * nk[n];
      nyq[n] = emul(n, PT, u0n[n]) + co(n, 3, PT) * (2.f * nk[n] - n0[n]);
    }
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, Nk, nk, ln, lds, rt, a);  // N2
    EXMI_EACH(S[n][s] = A[n][s] + co(n, 6, s) * Nk[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) nyq[n] = an[n] + co(n, 6, PT) * nk[n];
  } else {
    cf N0[NT][PT], Nk[NT][PT], A[NT][PT], U0[NT][PT], S1[NT][PT];
    cf n0[NT], nk[NT], an[NT], u0n[NT], s1n[NT];
    EXMI_EACH(U0[n][s] = S[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) u0n[n] = nyq[n];
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, N0, n0, ln, lds, rt, a);
    EXMI_EACH(A[n][s] = emul(n, s, U0[n][s]) + co(n, 3, s) * N0[n][s];
              S1[n][s] = hmul(n, s, U0[n][s]) + co(n, 2, s) * N0[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      an[n] = emul(n, PT, u0n[n]) + co(n, 3, PT) * n0[n];
      s1n[n] = hmul(n, PT, u0n[n]) + co(n, 2, PT) * n0[n];
    }
    nl_eval<KIND, M, PT, NT, LAY>(S1, s1n, Nk, nk, ln, lds, rt, a);  // N1
    EXMI_EACH(A[n][s] = A[n][s] + (co(n, 4, s) * 2.f) * Nk[n][s];
              S[n][s] = hmul(n, s, U0[n][s]) + co(n, 2, s) * Nk[n][s]);  // stage 2
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      an[n] = an[n] + (co(n, 4, PT) * 2.f) * nk[n];
      nyq[n] = hmul(n, PT, u0n[n]) + co(n, 2, PT) * nk[n];
    }
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, Nk, nk, ln, lds, rt, a);  // N2
    EXMI_EACH(A[n][s] = A[n][s] + (co(n, 4, s) * 2.f) * Nk[n][s];
              S[n][s] = hmul(n, s, S1[n][s]) + co(n, 2, s) * (2.f * Nk[n][s] - N0[n][s]));  // stage 3
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      an[n] = an[n] + (co(n, 4, PT) * 2.f) * nk[n];
      nyq[n] = hmul(n, PT, s1n[n]) + co(n, 2, PT) * (2.f * nk[n] - n0[n]);
    }
    nl_eval<KIND, M, PT, NT, LAY>(S, nyq, Nk, nk, ln, lds, rt, a);  // N3
    EXMI_EACH(S[n][s] = A[n][s] + co(n, 5, s) * Nk[n][s]);
#pragma unroll
    for (int n = 0; n < NT; ++n) nyq[n] = an[n] + co(n, 5, PT) * nk[n];
  }
#undef EXMI_EACH
}

// ------------------------------------------- folded ETDRK2 (constant cutoff)
// For the convection and gradient-norm kinds the nonlinear term is a fixed
// per-mode operator applied to the forward transform X of a product:
// 2 N_hat = f(k) m(k) op(X) with op(X) = i X, f = c k (conservative
// convection) or op(X) = X, f = c (non-conservative convection in its exact
// form; gradient norm, f(0) = 0 with the zero-mode fix), m the dealias mask.
// The rollout kernels with a compile-time cutoff fold f m into the ETDRK2
// coefficients when they stage the tables, so a nonlinear evaluation ends at
// the forward transform. With sc = 1/(2N) (physical rollouts; sc = 1 for the
// Fourier-resident one) they also fold the output scale in:
//   S1 = (sc E) S + (sc c1 f m) op(X0)                 = sc (E S + c1 N0)
//   S' = S1 + (sc c2 f m) op(K X1 - X0),  K = 1/sc^2   = sc S_next
// X1, the transform of the product of the scaled stage state, carries sc^2,
// and K X1 - X0 is one fused multiply-add with an exact power-of-two K. The
// inverse transform of S' is then the new physical state itself. Only the
// rounding of the per-mode table products changes (etdrk/_etdrk_2.py:91-102).
// The tables stay three floats per mode, as the unfolded ones: the launch
// fills every CU's LDS at the batch sizes of interest.
template <int KIND>
constexpr bool folded_kind() {
  return KIND == NL_CONV_CONS || KIND == NL_CONV_NONCONS || KIND == NL_GRADNORM;
}

// X: forward transform of the nonlinear product of 2 u_hat (paired layout),
// modes above KC left unmasked (their folded coefficients are zero).
template <int KIND, int M, int PT, int NT, int KC>
__device__ __forceinline__ void nl_product(const cf (*S)[PT], const cf* nyq, cf (*X)[PT], cf* xn,
                                           const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& rt, float kscale) {
  const int j = ln.j;
  if constexpr (KIND == NL_CONV_CONS) {  // u^2
#pragma unroll
    for (int n = 0; n < NT; ++n) masked_copy<M, PT, PAIRED>(S[n], nyq[n], X[n], xn[n], j, KC);
    irfft_paired<M, PT, NT>(X, xn, ln, lds, rt, KC);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) X[n][s] = cf{X[n][s].re * X[n][s].re, X[n][s].im * X[n][s].im};
  } else if constexpr (KIND == NL_CONV_NONCONS) {  // u u_x
    cf V[NT][PT];
    cf vn[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      masked_copy<M, PT, PAIRED>(S[n], nyq[n], X[n], xn[n], j, KC);
      masked_deriv<M, PT, PAIRED>(S[n], nyq[n], V[n], vn[n], j, KC, kscale);
    }
    irfft_paired<M, PT, NT>(X, xn, ln, lds, rt, KC);
    irfft_paired<M, PT, NT>(V, vn, ln, lds, rt, KC);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) X[n][s] = cf{X[n][s].re * V[n][s].re, X[n][s].im * V[n][s].im};
  } else {  // (u_x)^2
#pragma unroll
    for (int n = 0; n < NT; ++n) masked_deriv<M, PT, PAIRED>(S[n], nyq[n], X[n], xn[n], j, KC, kscale);
    irfft_paired<M, PT, NT>(X, xn, ln, lds, rt, KC);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int s = 0; s < PT; ++s) X[n][s] = cf{X[n][s].re * X[n][s].re, X[n][s].im * X[n][s].im};
  }
  rfft_paired<M, PT, NT>(X, xn, ln, lds, rt, KC);
}

// Folded coefficients of one mode k: co = [sc E, sc c1 f m, sc c2 f m].
template <int KIND, int KC>
__device__ __forceinline__ void fold_coefs(int k, float E, float c1, float c2, float cop, float sc, int zero_mode_fix,
                                           float* co) {
  float f = KIND == NL_CONV_CONS ? cop * (float)k : cop;
  if (KIND == NL_GRADNORM && k == 0 && zero_mode_fix) f = 0.f;
  if (k > KC) f = 0.f;
  co[0] = sc * E;
  co[1] = (sc * c1) * f;
  co[2] = (sc * c2) * f;
}

// Compact staging of the folded tables: sc E for every slot, the two
// nonlinear coefficients only for slots not wholly above the cutoff (they
// are zero there), lane-private in LDS; the mode M/2 (the extra register,
// meaningful in lane 0) keeps its three in registers. co(n, i, s) as
// CoefLds.
template <int M, int PT, int NT, int KC>
struct FoldTables {
  using MD = Modes<M, PT, PAIRED>;
  static constexpr int T = M / PT;
  __host__ __device__ static constexpr int kept_before(int s) {
    return s == 0 ? 0 : kept_before(s - 1) + (MD::klo(s - 1) <= KC ? 1 : 0);
  }
  static constexpr int NK = kept_before(PT);
  static constexpr int PER = PT + 2 * NK;  // floats per lane and transform
  static constexpr int FLOATS = PER * T;   // per transform
  float* base;
  int j;
  float ex[NT][3];
  __device__ __forceinline__ int at(int n, int i, int s) const {
    const int row = i == 0 ? s : PT + (i - 1) * NK + kept_before(s);
    return (n * PER + row) * T + j;
  }
  __device__ __forceinline__ float operator()(int n, int i, int s) const {
    return s == PT ? ex[n][i] : base[at(n, i, s)];
  }
  __device__ __forceinline__ void put(int n, int i, int s, float x) {
    if (s == PT) ex[n][i] = x;
    else if (i == 0 || MD::klo(s) <= KC) base[at(n, i, s)] = x;
  }
};

// K = 1/sc^2 (1 for the Fourier-resident rollout)
template <int KIND, int M, int PT, int NT, int KC, class Coef>
__device__ __forceinline__ void etdrk2_folded(float K, cf (*S)[PT], cf* nyq, const Coef& co, const Lane<M, PT>& ln, cf* lds,
                                              const TwRegs<M, PT>& rt, float kscale) {
  using MD = Modes<M, PT, PAIRED>;
  static_assert(MD::kextra <= KC, "the mode M/2 is kept by every 2/3 cutoff");
  // acc + g op(x) in two fused multiply-adds
  auto fop = [](cf acc, float g, cf x) -> cf {
    return KIND == NL_CONV_CONS ? cf{fmaf(-g, x.im, acc.re), fmaf(g, x.re, acc.im)}
                                : cf{fmaf(g, x.re, acc.re), fmaf(g, x.im, acc.im)};
  };
  cf X0[NT][PT], X1[NT][PT];
  cf x0n[NT], x1n[NT];
  nl_product<KIND, M, PT, NT, KC>(S, nyq, X0, x0n, ln, lds, rt, kscale);
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf es = co(n, 0, s) * S[n][s];
      S[n][s] = MD::klo(s) > KC ? es : fop(es, co(n, 1, s), X0[n][s]);
    }
    nyq[n] = fop(co(n, 0, PT) * nyq[n], co(n, 1, PT), x0n[n]);
  }
  nl_product<KIND, M, PT, NT, KC>(S, nyq, X1, x1n, ln, lds, rt, kscale);
  auto kdiff = [K](cf x1, cf x0) { return cf{fmaf(K, x1.re, -x0.re), fmaf(K, x1.im, -x0.im)}; };
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int s = 0; s < PT; ++s)
      if (MD::klo(s) <= KC) S[n][s] = fop(S[n][s], co(n, 2, s), kdiff(X1[n][s], X0[n][s]));
    nyq[n] = fop(nyq[n], co(n, 2, PT), kdiff(x1n[n], x0n[n]));
  }
}

// ------------------------------------------------------------- indexing
// Transform n of a workgroup: G groups of T lanes, each group NT transforms.
template <int M, int PT, int NT>
struct Slots {
  static constexpr int T = M / PT;
  static constexpr int WG = T > 64 ? T : 64;
  static constexpr int G = WG / T;
  static constexpr int PER_BLOCK = G * NT;
};

// --------------------------------------------------------------- kernels
#ifndef EXMI_STEP_WAVES_PER_EU
#define EXMI_STEP_WAVES_PER_EU 1
#endif
// Spectral layout of the step and rollout kernels (fft1d.hpp).
#ifndef EXMI_LAYOUT_1D
#define EXMI_LAYOUT_1D PAIRED
#endif

// Dealias cutoff of the default dealiasing_fraction 2/3 for N = 2M points:
// floor(2/3 * M - 1) (nonlin_fun/_base.py:59-71; never an integer boundary
// for power-of-two M). Kernels instantiated with KC >= 0 take the cutoff as a
// compile-time constant, so the per-slot mask classes resolve statically.
constexpr int default_kmax(int M) { return (2 * M - 3) / 3; }

// One step (or one nonlinear evaluation) of every transform of the batch.
// Tables are staged through lane-private LDS to keep register pressure low.
template <int M, int PT, int KIND, int ORDER, int LAY = EXMI_LAYOUT_1D, bool CPLX = false>
__global__ void __launch_bounds__((M / PT) > 64 ? (M / PT) : 64)
    __attribute__((amdgpu_waves_per_eu(EXMI_STEP_WAVES_PER_EU))) step1d_kernel(Step1DArgs a) {
  constexpr int T = M / PT;
  using MD = Modes<M, PT, LAY>;
  constexpr int G = Slots<M, PT, 1>::G;
  constexpr int LDSN = LdsSize<M>::value;
  constexpr int NR = M + 1;
  constexpr int NCO = ncoef_of(ORDER, CPLX);
  using CL = CoefLds<M, PT, NCO>;
  constexpr float NN = 2.f * M;
  __shared__ cf lds_all[G * LDSN];
  __shared__ float lds_coef[G * CL::FLOATS];

  const int g = threadIdx.x / T;
  const int j = threadIdx.x % T;
  const int t_raw = blockIdx.x * G + g;
  const bool active = t_raw < a.ntrans;
  const int t = active ? t_raw : 0;
  cf* lds = lds_all + g * LDSN;
  const CL cl{lds_coef + g * CL::FLOATS, j};
  const int b = t / a.channels, c = t % a.channels;

  const Lane<M, PT> ln(j);
  TwRegs<M, PT> rt;
  rt.load(j, RealTables{a.tw, a.fw});
  cf S[1][PT];
  cf nyq[1] = {cf{0.f, 0.f}};

  // ---- load the state (as 2 u_hat when spectral)
  if (a.in_fourier) {
    const cf* in = reinterpret_cast<const cf*>(a.in) + (long long)b * a.in_bstride + (long long)c * NR;
#pragma unroll
    for (int s = 0; s < PT; ++s) S[0][s] = 2.f * in[MD::k(j, s)];
    nyq[0] = 2.f * in[MD::kextra];
  } else {
    const float2* in = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(a.in) +
                                                       (long long)b * a.in_bstride + (long long)c * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      float2 v = in[j + s * T];
      S[0][s] = cf{v.x, v.y};
    }
  }

  // ---- operator tables: HBM -> registers -> lane-private LDS
  if (!a.nonlin_only) {
    const float* tb = a.tables + ((long long)(a.lin_batch > 1 ? b : 0) * a.lin_channels +
                                  (a.lin_channels > 1 ? c : 0)) * (long long)NCO * NR;
    CoefRegs<NCO, PT, 1> cr;
    cr.template load<M, LAY>(0, tb, j);
    if (!a.in_fourier) rfft_l<LAY, M, PT, 1>(S, nyq, ln, lds, rt);
#pragma unroll
    for (int i = 0; i < NCO; ++i)
#pragma unroll
      for (int s = 0; s <= PT; ++s) cl.put(0, i, s, cr(0, i, s));
  } else if (!a.in_fourier) {
    rfft_l<LAY, M, PT, 1>(S, nyq, ln, lds, rt);
  }

  if (a.nonlin_only) {
    cf Nn[1][PT];
    cf nn[1];
    nl_eval<KIND, M, PT, 1, LAY>(S, nyq, Nn, nn, ln, lds, rt, a);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[0][s] = Nn[0][s];
    nyq[0] = nn[0];
  } else {
    etdrk_step<KIND, ORDER, M, PT, 1, CL, LAY, CPLX>(S, nyq, cl, ln, lds, rt, a);
  }

  // ---- store (2 u_hat -> u_hat, or N * 2 u -> u)
  if (a.out_fourier) {
    if (active) {
      cf* out = reinterpret_cast<cf*>(a.out) + (long long)b * a.out_bstride + (long long)c * NR;
#pragma unroll
      for (int s = 0; s < PT; ++s) out[MD::k(j, s)] = 0.5f * S[0][s];
      if (j == 0) out[MD::kextra] = 0.5f * nyq[0];
    }
  } else {
    irfft_l<LAY, M, PT, 1>(S, nyq, ln, lds, rt);
    if (active) {
      constexpr float r2n = 1.f / (2.f * NN);
      float2* out = reinterpret_cast<float2*>(reinterpret_cast<float*>(a.out) + (long long)b * a.out_bstride +
                                              (long long)c * 2 * M);
#pragma unroll
      for (int s = 0; s < PT; ++s) out[j + s * T] = make_float2(S[0][s].re * r2n, S[0][s].im * r2n);
    }
  }
}

template <int M, int PT, int KIND, int ORDER, int LAY = EXMI_LAYOUT_1D, bool CPLX = false>
inline hipError_t launch_step1d(const Step1DArgs& a, hipStream_t st) {
  constexpr int WG = Slots<M, PT, 1>::WG;
  constexpr int G = Slots<M, PT, 1>::G;
  const int grid = (a.ntrans + G - 1) / G;
  hipLaunchKernelGGL((step1d_kernel<M, PT, KIND, ORDER, LAY, CPLX>), dim3(grid), dim3(WG), 0, st, a);
  return hipGetLastError();
}

// Multi-step rollout: each group of lanes keeps NT transforms' states,
// operator tables and twiddles on chip for `nsteps` steps. Per step the
// arithmetic is exactly that of step1d_kernel (physical state in float32 ->
// forward transform -> ETDRK step -> inverse transform -> float32 state);
// the new state is written to traj[b, t0 + step] (or only the final state
// when `final_only`). Every wave runs the same bounded loop; there is no
// inter-wave communication.
struct Roll1DArgs {
  Step1DArgs s;            // in, in_bstride: initial states (physical)
  float* traj;             // output base
  long long traj_bstride;  // floats between trajectories
  long long traj_tstride;  // floats between consecutive outputs of one trajectory
  int nsteps;
  int final_only;          // 1: write only the last state (traj_tstride unused)
  int* first_bad = nullptr;  // failure detection, or nullptr: per trajectory, atomicMin of the first
                             // written slot (bad_offset + step) holding a value that is not finite or
                             // exceeds `bound` in magnitude
  float bound = 0.f;
  int bad_offset = 0;
};

// Write the NT physical states (N x, times `scale`) of a step to the
// trajectory buffer.
template <int M, int PT, int NT>
__device__ __forceinline__ void store_states(const cf (*S)[PT], float scale, const Roll1DArgs& ra, const int* tb_,
                                             const int* cb_, const bool* act, int step, int j) {
  constexpr int T = M / PT;
  const long long tofs = ra.final_only ? 0 : (long long)step * ra.traj_tstride;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if (act[n]) {
      float2* out = reinterpret_cast<float2*>(ra.traj + (long long)tb_[n] * ra.traj_bstride + tofs +
                                              (long long)cb_[n] * 2 * M);
#pragma unroll
      for (int s = 0; s < PT; ++s) out[j + s * T] = make_float2(S[n][s].re * scale, S[n][s].im * scale);
    }
  }
}

// FR (Fourier-resident): the spectral state stays on chip from step to step
// (one forward transform before the loop); each step's physical state is
// the inverse transform of a copy. It differs from the physical path only in
// rounding (no float32 round trip of the state through physical space), and
// it saves one forward transform per step.
template <int M, int PT, int KIND, int ORDER, int NT, bool FR = false, int KC = -1, int LAY = EXMI_LAYOUT_1D,
          bool CPLX = false>
__global__ void __launch_bounds__((M / PT) > 64 ? (M / PT) : 64) roll1d_kernel(Roll1DArgs ra) {
  Step1DArgs a = ra.s;
  if constexpr (KC >= 0) a.kmax = KC;  // the caller launches KC only when it equals the plan's cutoff
  constexpr int T = M / PT;
  using MD = Modes<M, PT, LAY>;
  constexpr int G = Slots<M, PT, NT>::G;
  constexpr int LDSN = LdsSize<M>::value;
  constexpr int NR = M + 1;
  constexpr int NCO = ncoef_of(ORDER, CPLX);
  constexpr bool FOLD = KC >= 0 && ORDER == 2 && LAY == PAIRED && !CPLX && folded_kind<KIND>();
  using CL = std::conditional_t<FOLD, FoldTables<M, PT, NT, (KC < 0 ? 0 : KC)>, CoefLds<M, PT, NCO>>;
  constexpr float NN = 2.f * M;
  constexpr float r2n = 1.f / (2.f * NN);
  __shared__ cf lds_all[G * LDSN];
  __shared__ float lds_coef[G * NT * CL::FLOATS];

  const int g = threadIdx.x / T;
  const int j = threadIdx.x % T;
  cf* lds = lds_all + g * LDSN;
  CL cl{lds_coef + g * NT * CL::FLOATS, j};
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> rt;
  rt.load(j, RealTables{a.tw, a.fw});

  int tb_[NT], cb_[NT];
  bool act[NT];
  cf S[NT][PT];
  cf nyq[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const int t_raw = (blockIdx.x * G + g) * NT + n;
    act[n] = t_raw < a.ntrans;
    const int t = act[n] ? t_raw : 0;
    tb_[n] = t / a.channels;
    cb_[n] = t % a.channels;
    const float2* in = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(a.in) +
                                                       (long long)tb_[n] * a.in_bstride + (long long)cb_[n] * 2 * M);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      float2 v = in[j + s * T];
      S[n][s] = cf{v.x, v.y};
    }
    const float* tb = a.tables + ((long long)(a.lin_batch > 1 ? tb_[n] : 0) * a.lin_channels +
                                  (a.lin_channels > 1 ? cb_[n] : 0)) * (long long)NCO * NR;
    if constexpr (FOLD) {
      constexpr float q2 = 2.f / (8.f * NN * NN);
      const float cop = KIND == NL_CONV_CONS ? -a.scale * 0.5f * a.kscale * q2
                                             : (KIND == NL_GRADNORM ? -a.scale * 0.5f * q2 : -a.scale * q2);
      const float sc = FR ? 1.f : r2n;
#pragma unroll
      for (int s = 0; s <= PT; ++s) {
        const int k = s < PT ? MD::k(j, s) : MD::kextra;
        float co[3];
        fold_coefs<KIND, KC>(k, tb[k], tb[NR + k], tb[2 * NR + k], cop, sc, a.zero_mode_fix, co);
#pragma unroll
        for (int i = 0; i < 3; ++i) cl.put(n, i, s, co[i]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < NCO; ++i) {
#pragma unroll
        for (int s = 0; s < PT; ++s) cl.put(n, i, s, tb[i * NR + MD::k(j, s)]);
        cl.put(n, i, PT, tb[i * NR + MD::kextra]);
      }
    }
  }

  // failure detection: the first step whose written state fails the bound
  // (a lane's own values; the trajectory's minimum is taken at the end)
  int fb[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) fb[n] = INT_MAX;
  auto check = [&](const cf (*V)[PT], float scale, int step) {
    if (!ra.first_bad) return;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      bool bad = false;
#pragma unroll
      for (int s = 0; s < PT; ++s)
        bad |= !(fabsf(V[n][s].re * scale) <= ra.bound) | !(fabsf(V[n][s].im * scale) <= ra.bound);
      if (bad && fb[n] == INT_MAX) fb[n] = step;
    }
  };

  if constexpr (FR) rfft_l<LAY, M, PT, NT>(S, nyq, ln, lds, rt);
  for (int step = 0; step < ra.nsteps; ++step) {
    if constexpr (!FR) rfft_l<LAY, M, PT, NT>(S, nyq, ln, lds, rt);
    if constexpr (FOLD) etdrk2_folded<KIND, M, PT, NT, KC>(FR ? 1.f : 1.f / (r2n * r2n), S, nyq, cl, ln, lds, rt, a.kscale);
    else etdrk_step<KIND, ORDER, M, PT, NT, CL, LAY, CPLX>(S, nyq, cl, ln, lds, rt, a);
    const bool emit = !ra.final_only || step == ra.nsteps - 1;
    if constexpr (FR) {
      if (emit) {
        cf O[NT][PT];
        cf on[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          on[n] = nyq[n];
#pragma unroll
          for (int s = 0; s < PT; ++s) O[n][s] = S[n][s];
        }
        irfft_l<LAY, M, PT, NT>(O, on, ln, lds, rt);
        store_states<M, PT, NT>(O, r2n, ra, tb_, cb_, act, step, j);
        check(O, r2n, step);
      }
    } else {
      irfft_l<LAY, M, PT, NT>(S, nyq, ln, lds, rt);
      // back to the float32 physical state, exactly as a stored-and-reloaded
      // state of the single-step path (the folded step's state carries the
      // scale already)
      if constexpr (!FOLD) {
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
          for (int s = 0; s < PT; ++s) S[n][s] = cf{S[n][s].re * r2n, S[n][s].im * r2n};
      }
      if (emit) {
        store_states<M, PT, NT>(S, 1.f, ra, tb_, cb_, act, step, j);
        check(S, 1.f, step);
      }
    }
  }
  if (ra.first_bad) {
#pragma unroll
    for (int n = 0; n < NT; ++n)
      if (act[n] && fb[n] != INT_MAX) atomicMin(ra.first_bad + tb_[n], ra.bad_offset + (ra.final_only ? 0 : fb[n]));
  }
}

template <int M, int PT, int KIND, int ORDER, int NT, bool FR = false, int KC = -1, int LAY = EXMI_LAYOUT_1D,
          bool CPLX = false>
inline hipError_t launch_roll1d(const Roll1DArgs& ra, hipStream_t st) {
  constexpr int WG = Slots<M, PT, NT>::WG;
  constexpr int PB = Slots<M, PT, NT>::PER_BLOCK;
  const int grid = (ra.s.ntrans + PB - 1) / PB;
  if (KC >= 0 && ra.s.kmax != KC) return hipErrorInvalidValue;
  hipLaunchKernelGGL((roll1d_kernel<M, PT, KIND, ORDER, NT, FR, KC, LAY, CPLX>), dim3(grid), dim3(WG), 0, st, ra);
  return hipGetLastError();
}

}  // namespace exmi
