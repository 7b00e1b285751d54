// Workgroup-resident 2-D ETDRK2 rollout for N x N grids, N <= 128 (gfx950).
//
// One workgroup owns one instance (one single-channel trajectory) for all
// steps of a rollout: its spectral state stays in registers, the
// transposes between the column (x) and row (y) transforms go through LDS,
// and the only HBM traffic per step is the state written to the trajectory
// (plus reads of the operator tables, served from the caches). The step is
// the one of etdrk/_etdrk_2.py:91-102 with the nonlinear function
// evaluated as nonlin_fun/_base.py:99-137 prescribes (dealias before every
// inverse transform, after every forward one):
//   gradient norm     -b/2 F[(d_x u)^2 + (d_y u)^2]      (_gradient_norm.py:84-101)
//   vorticity         -b F[u d_x w + v d_y w] + forcing   (_vorticity_convection.py:78-99, 166-182)
//
// The spectrum is carried from step to step and each step's physical state
// is its inverse transform: this kernel serves plans for which the forward
// transform of an emitted state returns the spectrum exactly in exact
// arithmetic (a real L; the dealias mask removes the Nyquist row and column
// from the nonlinear term), the reference's rollout (_utils.py:167-172)
// then being the same function up to float32 rounding.
//
// Layouts inside the workgroup (PT = 8 points per lane everywhere):
//   columns: N/2 groups of TC = N/8 lanes; group g owns ky = g, lane j the
//            entries kx-index j + TC s (s < 8) in registers; group 0 also
//            owns ky = N/2 (the Nyquist column, kept in LDS `sb`), which the
//            nonlinear term never reaches (dealiased) and which its inverse
//            transform packs with column 0: both are Hermitian in kx, so one
//            complex transform of A + iB returns ifft(A) and ifft(B) as its
//            real and imaginary parts.
//   rows:    N groups of TR = N/16 lanes; group x owns the row u(x, :) as
//            N/2 complex samples u(x, 2n) + i u(x, 2n + 1), n = j + TR s.
//   LDS:     the transforms' exchange regions (column and row transforms
//            share them, separated by barriers), then one [ky][x] complex
//            buffer with the padded stride LD = N + 4 for the field being
//            transposed (a column group touches only its row ky of it, a row
//            group only its column x); the full spectrum of an emitted state
//            and the first nonlinear term share the space after the exchange
//            regions, as they are never live together.
//
// Scaling as in the multi-pass 2-D kernels: forward transforms give 2 X,
// inverse ones N^2 x; the state is carried as 2 u_hat and the nonlinear
// terms as 2 N_hat.
#pragma once
#include "step2d.hpp"

namespace exmi {

struct Roll2DArgs {
  Step2DArgs s;            // tables, twiddles, kscale, kmax, scale, zero_mode_fix, forcing
  const float* u0;         // (ninst, N, N) initial states, stride u0_bstride floats
  long long u0_bstride;
  float* traj;             // output base
  long long traj_bstride;  // floats between instances
  long long traj_tstride;  // floats between consecutive outputs of one instance
  int nsteps;
  int final_only;          // 1: write only the last state
  int ninst;
};

template <int N>
struct R2Cfg {
  static constexpr int M = N / 2;  // complex length of a row transform
  static constexpr int NY = N / 2 + 1;
  static constexpr int PT = 8;
  static constexpr int TC = N / PT;  // lanes per column transform
  static constexpr int GC = N / 2;   // column groups
  static constexpr int TR = M / PT;  // lanes per row transform
  static constexpr int WG = GC * TC;
  static constexpr int LD = N + 4;   // padded [ky][x] row
  static constexpr int LDSC = LdsSize<N>::value;
  static constexpr int LDSR = LdsSize<M>::value;
  static constexpr int mx(int a, int b) { return a > b ? a : b; }
  // kept ky columns of the default 2/3 dealias cutoff, floor(2/3 (N/2) - 1)
  // + 1 (nonlin_fun/_base.py:59-71): the host launches this kernel only for
  // cutoffs up to it
  static constexpr int KC = (2 * M - 3) / 3 + 1;
  static constexpr int SCR = mx(GC * LDSC, N * LDSR);  // exchange regions of every column or row transform
  static constexpr int FB = SCR;                       // field buffer [ky < KC][x]
  static constexpr int N0 = FB + KC * LD;              // first nonlinear term [ky < KC][kx]
  // the full spectrum [ky <= N/2][x] of an emitted state overlays FB and N0
  static constexpr int ARENA = mx(N0 + KC * N, FB + NY * LD);
  static_assert(WG == N * TR, "one column lane per row lane");
  static_assert(TC <= 64 && TR <= 64 && TR >= 1, "transforms within one wavefront");
};

template <int KIND>
constexpr int roll2d_batches() {
  return KIND == NL_VORT ? 2 : 1;  // field pairs per nonlinear evaluation
}

template <int N, int KIND>
__global__ void __launch_bounds__(R2Cfg<N>::WG) roll2d_kernel(Roll2DArgs ra) {
  using C = R2Cfg<N>;
  constexpr int M = C::M, NY = C::NY, PT = C::PT, TC = C::TC, TR = C::TR, LD = C::LD;
  constexpr float NN = (float)N;
  constexpr float q2 = 2.f / (8.f * NN * NN * NN * NN);  // 2 N_hat from 8 N^4 F[product]
  constexpr float rn2 = 1.f / (2.f * NN * NN);           // u from the inverse of 2 u_hat
  static_assert(KIND == NL_GRADNORM || KIND == NL_VORT, "workgroup-resident kinds");
  __shared__ cf arena[C::ARENA];
  __shared__ cf sb[N];  // 2 u_hat of column ky = N/2, by kx index
  cf* const fb = arena + C::FB;
  cf* const n0s = arena + C::N0;

  const Step2DArgs& a = ra.s;
  const int inst = blockIdx.x;
  const int tid = threadIdx.x;
  const int g = tid / TC, jc = tid % TC;   // column group (ky = g) and lane
  const int xr = tid / TR, jr = tid % TR;  // row x and lane
  const int ky = g;
  const int kc = a.kmax + 1 < C::KC ? a.kmax + 1 : C::KC;  // the host launches this kernel only for kmax < KC
  const long long full = (long long)NY * N;
  const float* tb = a.tables + (long long)(a.lin_batch > 1 ? inst : 0) * a.ncoef * full;
  const float kap_y = a.kscale * (float)ky;
  // signed kx of slot s and the dealias predicate |kx| <= kmax (recomputed
  // where used: registers are the scarce resource here)
  auto kx_of = [&](int s) {
    const int idx = jc + TC * s;
    return idx < N / 2 ? idx : idx - N;
  };
  auto keep_x = [&](int s) {
    const int kx = kx_of(s);
    return (kx < 0 ? -kx : kx) <= a.kmax;
  };
  const bool kept_col = ky < kc;
  const Lane<N, PT> lnc(jc);
  const Lane<M, PT> lnr(jr);
  cf* ldc = arena + g * C::LDSC;   // column transform exchange region
  cf* ldr = arena + xr * C::LDSR;  // row transform exchange region
  auto coef = [&](int i, int kyy, int s) { return tb[(long long)i * full + (long long)kyy * N + jc + TC * s]; };

  // ---- nonlinear term 2 N_hat of the state Sx (2 u_hat, this group's
  // column) into Nh; every thread of the workgroup takes part. The fields
  // pass one at a time: column inverse transform into the field buffer,
  // then the rows' inverse transforms and the product accumulated in
  // registers.
  auto nonlinear = [&](const cf* Sx, cf* Nh) {
    constexpr int NFLD = KIND == NL_VORT ? 4 : 2;
    cf P[PT], A[PT];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) {
      if (kept_col) {
        cf v[1][PT];
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf w = keep_x(s) ? Sx[s] : cf{0.f, 0.f};
          const float kx = a.kscale * (float)kx_of(s);
          float k = kx;  // the field is i k w or i k psi
          cf base = w;
          if constexpr (KIND == NL_GRADNORM) {
            k = f == 0 ? kx : kap_y;  // (d_x u, d_y u)
          } else {
            // psi = Laplace^-1 w (1 at the mean mode); fields in order
            // u = i ky psi, w_x = i kx w, v = -i kx psi, w_y = i ky w
            const float lap = -(kx * kx + kap_y * kap_y);
            const float inv = lap == 0.f ? 1.f : 1.f / lap;
            if (f == 0 || f == 2) base = inv * w;
            k = f == 0 ? kap_y : f == 1 ? kx : f == 2 ? -kx : kap_y;
          }
          v[0][s] = cf{-k * base.im, k * base.re};
        }
        TwRegs<N, PT> trc;
        trc.template load_pass<1>(jc, a.tw_col);
        cfft<N, PT, true, 1>(v, lnc, ldc, trc);
#pragma unroll
        for (int s = 0; s < PT; ++s) fb[ky * LD + jc + TC * s] = v[0][s];
      }
      __syncthreads();
      {
        cf z[1][PT];
        cf nyq[1] = {cf{0.f, 0.f}};  // the Nyquist column is dealiased
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const int k = jr + TR * s;
          z[0][s] = k < kc ? fb[k * LD + xr] : cf{0.f, 0.f};
        }
        TwRegs<M, PT> trr;
        trr.load(jr, RealTables{a.tw_row, a.fw_row});
        irfft<M, PT, 1>(z, nyq, lnr, ldr, trr);
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          if constexpr (KIND == NL_GRADNORM) {
            P[s] = f == 0 ? cf{q.re * q.re, q.im * q.im} : cf{fmaf(q.re, q.re, P[s].re), fmaf(q.im, q.im, P[s].im)};
          } else {
            if (f == 0 || f == 2) A[s] = q;
            if (f == 1) P[s] = cf{A[s].re * q.re, A[s].im * q.im};
            if (f == 3) P[s] = cf{fmaf(A[s].re, q.re, P[s].re), fmaf(A[s].im, q.im, P[s].im)};
          }
        }
      }
      __syncthreads();  // the exchange regions are reused by the next column transforms
    }
    (void)A;
    // forward transform of the product: rows (modes above kmax not formed),
    // each row group writing its own column x of the field buffer
    {
      cf z[1][PT];
      cf nyq[1];
#pragma unroll
      for (int s = 0; s < PT; ++s) z[0][s] = P[s];
      TwRegs<M, PT> trr;
      trr.load(jr, RealTables{a.tw_row, a.fw_row});
      rfft<M, PT, 1>(z, nyq, lnr, ldr, trr, a.kmax);
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        const int k = jr + TR * s;
        if (k < kc) fb[k * LD + xr] = z[0][s];
      }
    }
    __syncthreads();
    // columns: forward transform along x, post-dealias, scale, forcing
    cf v[1][PT];
    if (kept_col) {
#pragma unroll
      for (int s = 0; s < PT; ++s) v[0][s] = fb[ky * LD + jc + TC * s];
      TwRegs<N, PT> trc;
      trc.template load_pass<1>(jc, a.tw_col);
      cfft<N, PT, false, 1>(v, lnc, ldc, trc);
    }
    const float cs = KIND == NL_VORT ? -a.scale * q2 : -a.scale * 0.5f * q2;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      cf r = (kept_col && keep_x(s)) ? cs * v[0][s] : cf{0.f, 0.f};
      if (KIND == NL_GRADNORM && a.zero_mode_fix && ky == 0 && jc + TC * s == 0) r = cf{0.f, 0.f};
      if (KIND == NL_VORT && a.inj_mode == ky && jc + TC * s == 0) r = r + cf{2.f * a.inj_value, 0.f};
      Nh[s] = r;
    }
    __syncthreads();  // the exchange regions are reused next
  };

  // ---- inverse transform of the state (2 u_hat) -> physical rows, written
  // to the trajectory at output slot t
  auto emit = [&](const cf* S, long long t) {
    cf v[1][PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      v[0][s] = S[s];
      if (g == 0) {  // A + i B with B = column ky = N/2
        const cf bq = sb[jc + TC * s];
        v[0][s] = cf{S[s].re - bq.im, S[s].im + bq.re};
      }
    }
    {
      TwRegs<N, PT> trc;
      trc.template load_pass<1>(jc, a.tw_col);
      cfft<N, PT, true, 1>(v, lnc, ldc, trc);
    }
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const int x = jc + TC * s;
      fb[ky * LD + x] = v[0][s];  // ky = 0: the real part is ifft(A); irfft ignores the rest
      if (g == 0) fb[M * LD + x] = cf{v[0][s].im, 0.f};
    }
    __syncthreads();
    cf z[1][PT];
    cf nyq[1];
#pragma unroll
    for (int s = 0; s < PT; ++s) z[0][s] = fb[(jr + TR * s) * LD + xr];
    nyq[0] = fb[M * LD + xr];
    {
      TwRegs<M, PT> trr;
      trr.load(jr, RealTables{a.tw_row, a.fw_row});
      irfft<M, PT, 1>(z, nyq, lnr, ldr, trr);
    }
    float2* out = reinterpret_cast<float2*>(ra.traj + (long long)inst * ra.traj_bstride + t * ra.traj_tstride +
                                            (long long)xr * N);
#pragma unroll
    for (int s = 0; s < PT; ++s) out[jr + TR * s] = make_float2(z[0][s].re * rn2, z[0][s].im * rn2);
    __syncthreads();
  };

  // ---- initial spectrum: rows r2c, then every column (group 0: ky = 0 and
  // ky = N/2) c2c
  cf S[PT];
  {
    cf z[1][PT];
    cf nyq[1];
    const float2* u = reinterpret_cast<const float2*>(ra.u0 + (long long)inst * ra.u0_bstride + (long long)xr * N);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const float2 q = u[jr + TR * s];
      z[0][s] = cf{q.x, q.y};
    }
    TwRegs<M, PT> trr;
    trr.load(jr, RealTables{a.tw_row, a.fw_row});
    rfft<M, PT, 1>(z, nyq, lnr, ldr, trr);
#pragma unroll
    for (int s = 0; s < PT; ++s) fb[(jr + TR * s) * LD + xr] = z[0][s];
    if (jr == 0) fb[M * LD + xr] = nyq[0];
    __syncthreads();
    cf v[2][PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      v[0][s] = fb[ky * LD + jc + TC * s];
      v[1][s] = fb[(g == 0 ? M : ky) * LD + jc + TC * s];
    }
    TwRegs<N, PT> trc;
    trc.template load_pass<1>(jc, a.tw_col);
    cfft<N, PT, false, 2>(v, lnc, ldc, trc);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = v[0][s];
    if (g == 0) {
#pragma unroll
      for (int s = 0; s < PT; ++s) sb[jc + TC * s] = v[1][s];
    }
    __syncthreads();
  }

  // the forcing part of a nonlinear term outside the kept columns (zero
  // for the gradient norm)
  auto n_outside = [&](int s) -> cf {
    return (KIND == NL_VORT && a.inj_mode == ky && jc + TC * s == 0) ? cf{2.f * a.inj_value, 0.f} : cf{0.f, 0.f};
  };
  for (int step = 0; step < ra.nsteps; ++step) {
    {
      cf N0[PT];
      nonlinear(S, N0);
      // S1 = E S + c1 N0 (etdrk/_etdrk_2.py:95-96); N0 of the kept columns
      // waits in LDS; the Nyquist column has no nonlinear term: E S there
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        S[s] = coef(0, ky, s) * S[s] + coef(1, ky, s) * N0[s];
        if (kept_col) n0s[ky * N + jc + TC * s] = N0[s];
      }
    }
    if (g == 0) {
#pragma unroll
      for (int s = 0; s < PT; ++s) sb[jc + TC * s] = coef(0, M, s) * sb[jc + TC * s];
    }
    cf N1[PT];
    nonlinear(S, N1);
    // S' = S1 + c2 (N1 - N0) (:97-101)
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf n0 = kept_col ? n0s[ky * N + jc + TC * s] : n_outside(s);
      S[s] = S[s] + coef(2, ky, s) * (N1[s] - n0);
    }
    if (!ra.final_only) emit(S, step);
    else if (step == ra.nsteps - 1) emit(S, 0);
  }
}

template <int N, int KIND>
inline hipError_t launch_roll2d(const Roll2DArgs& ra, hipStream_t st) {
  using C = R2Cfg<N>;
  if (ra.s.kmax >= C::KC || ra.ninst < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL((roll2d_kernel<N, KIND>), dim3(ra.ninst), dim3(C::WG), 0, st, ra);
  return hipGetLastError();
}

}  // namespace exmi
