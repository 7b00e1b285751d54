// Fused 2-D ETDRK2 pseudo-spectral step for gfx950 (N x N periodic grid).
//
// A 2-D real transform is a real FFT along y (the last, contiguous axis) of
// every row followed by a complex FFT along x of every kept column, the
// reference's rfftn (_spectral.py:614-656) / irfftn (:659-721) order. One
// ETDRK2 step (etdrk/_etdrk_2.py:91-102) is a short chain of launches, each
// one pass over the field:
//
//   row RA : u rows -> r2c along y                        -> T [ky][x]
//   col CA : T column -> c2c along x = u_hat; keep S0; the nonlinear
//            evaluation's fields (dealiased derivatives of u_hat)
//            -> inverse c2c along x                       -> F fields [ky][x]
//   row RB : F fields -> c2r along y, nonlinear product, r2c -> P [ky][x]
//   col CB : P -> c2c = N_hat0 (post dealias, scale, forcing); S1 = E S0 +
//            c1 N0; fields of S1 -> inverse c2c          -> F fields
//   row RB : as above                                     -> P
//   col CC : P -> N_hat1; S' = S1 + c2 (N1 - N0) -> inverse c2c -> T
//   row RC : T -> c2r along y -> u' rows (in a rollout also r2c of u' again
//            -> T for the next step, i.e. RA fused)
//
// Row/column interface fields are column-major ([ky][x], one column
// contiguous) and transposed through an LDS tile in the row kernels, so every
// global access is coalesced. Only the ky <= kmax columns of the dealiased
// fields exist. Between launches the fields stay in HBM, largely served from
// the Infinity Cache at these sizes.
//
// Scaling as in 1-D: forward transforms yield 2 X, inverse ones N^2 x; the
// spectral state is carried as 2 u_hat and the nonlinear terms as 2 N_hat;
// the nonlinear coefficients absorb the powers of two (exact).
//
// Nonlinear functions: VorticityConvection2d(+Kolmogorov)
// (nonlin_fun/_vorticity_convection.py:65-99, 178-182),
// GradientNormNonlinearFun in 2-D (nonlin_fun/_gradient_norm.py:84-101),
// PolynomialNonlinearFun (nonlin_fun/_polynomial.py:64-76), the two-channel
// convections of a velocity field (nonlin_fun/_convection.py:104-171) and
// GeneralNonlinearFun (nonlin_fun/_general_nonlinear.py:89-119).
//
// Channels: every (trajectory, channel) pair is an instance of its own with
// its own spectral state; instance i is channel i % C of trajectory i / C.
// The convections couple the channels only through the row kernel's
// products, which read the velocity fields of all channels of the
// trajectory.
//
// The same row and column kernels also serve the standalone 2-D transforms:
// rfftn is RA then COL_FWD (column FFT written in the reference's [kx][ky]
// layout), irfftn is COL_INV (column inverse FFT from [kx][ky]) then RC.
#pragma once
#include "exmi_internal.hpp"
#include "fft1d.hpp"

// 4 points per lane up to these sizes, 8 above (build-time knobs; the
// defaults were the fastest measured for the mixed 128^2/256^2/512^2 batch)
#ifndef EXMI_ROW_PT4_MAX
#define EXMI_ROW_PT4_MAX 128
#endif
#ifndef EXMI_ROW_R
#define EXMI_ROW_R 16
#endif
#ifndef EXMI_COL_PT4_MAX
#define EXMI_COL_PT4_MAX 128
#endif

namespace exmi {

struct Step2DArgs {
  int channels;              // C: instances per trajectory
  long long tstride;         // complex entries of T per instance (>= Ny N and the products' NP Kc N)
  float g0, g1, g2;          // GeneralNonlinearFun scales, signs folded in
  int prod_nc;               // general kind: first product u (u_x + u_y) instead of u^2
                             // (single-channel non-conservative convection)
  const float* u_in;         // physical (batch, C, N, N) with a batch stride
  long long u_in_bstride;    // floats between instances
  float* u_out;              // physical output of RC
  long long u_out_bstride;
  cf* S0;                    // (ninst, Ny, N)   spectral state [ky][kx]
  cf* N0;                    // (ninst, Kc, N)   first nonlinear term, ky < Kc
  cf* S1;                    // (ninst, Ny, N)   stage state
  cf* F;                     // (ninst, NF, Kc, N) nonlinear fields [ky][x]
  cf* T;                     // (ninst, Ny, N)   row spectra / products [ky][x]
  cf* A;                     // (ninst, Ny, N)   ETDRK3/4 accumulator of the final combination
  const float* tables;       // (lin_batch, ncoef, Ny, N): ETDRK0 [E], ETDRK1 [E, c1], ETDRK2 [E, c1, c2], [ky][kx]
  int ncoef;
  int lin_batch;
  const cf* tw_row;          // row real FFT (M = N/2) tables
  const cf* fw_row;
  const cf* tw_col;          // column complex FFT (N points) twiddles
  float kscale;              // 2 pi / L
  int kmax;                  // dealias cutoff (|kx|, ky <= kmax)
  float scale;               // convection / gradient-norm scale b
  int zero_mode_fix;
  int inj_mode;              // Kolmogorov forcing wavenumber (ky index); -1: none
  float inj_value;           // forcing value at (kx = 0, ky = inj_mode), reference scaling
  int write_T;               // RC: also r2c the new state into T (rollout)
  int e_im;                  // coefficient index of Im E (complex L; Im E_half follows for orders 3-4); -1: real L
  int npoly;                 // polynomial coefficients c_0..c_{npoly-1}
  float poly[EXMI_MAX_POLY];
  int poly_lap;              // 1: scale * Laplace of the polynomial's transform (Cahn-Hilliard)
  int lin_channels;          // 1 or C: per-channel linear operators
  float feed, kill;          // Gray-Scott rates
  cf* spec;                  // standalone transforms: spectra (ninst, N, N/2+1), [kx][ky]
  cf* spec_out;              // exmi_nonlinear: N_hat (ninst, N, N/2+1), [kx][ky]
  long long spec_bstride;    // complex entries between instances
  int next_fields;           // COL_CCA: also form the fields of S' (the next step's first evaluation)
  float* Gout;               // gradient norm, RB/RCB: also store the physical gradient fields (the
                             // adjoint's G layout: (ninst, 2 stages, 2 fields, N, N)), or null
  int gstage;                // the stage those fields belong to (0: the state, 1: the stage state S1)
  int spec_io;               // exmi_step_fourier: COL_C0 reads u_hat from spec, and the last stage
                             // (C0, C1, CC, C3C, C4D) writes u_hat' to spec_out, both [kx][ky]
};

// CA/CB/CC: the ETDRK2 stages; CCA: CC that also keeps S' as the next
// step's S0 and forms its fields (the spectrum-resident chain, where the
// next step's forward transform of the emitted state is an exact identity
// and is skipped: CA of the next step fused into CC); C0: ETDRK0 (S' = E S); C1: ETDRK1's update
// (S' = E S0 + c1 N0, etdrk/_etdrk_1.py:77-82) after CA and RB.
// C3A..C3C: the three ETDRK3 stages after CA (etdrk/_etdrk_3.py:191-212),
// C4A..C4D: the four ETDRK4 stages after CA (etdrk/_etdrk_4.py:198-224); each
// consumes the product of the preceding RB and leaves the next stage's
// fields (or, last, the new state's inverse column transform in T).
enum Col2DMode {
  COL_CA = 0, COL_CB = 1, COL_CC = 2, COL_C0 = 3, COL_C1 = 4,
  COL_C3A = 5, COL_C3B = 6, COL_C3C = 7, COL_C4A = 8, COL_C4B = 9, COL_C4C = 10, COL_C4D = 11,
  COL_FWD = 12, COL_INV = 13,
  COL_CA_SPEC = 14, COL_NOUT = 15,
  COL_CCA = 16,
  COL_NUM_MODES = 17
};
// ROW_RCB: RC (the inverse row transforms of T into u_out) then RB (the
// products of the fields F into T), in one pass over the rows.
enum Row2DMode { ROW_RA = 0, ROW_RB = 1, ROW_RC = 2, ROW_RCB = 3, ROW_NUM_MODES = 4 };

// Nonlinear fields of one evaluation. Vorticity: (u, omega_x, v, omega_y),
// product u omega_x + v omega_y. Gradient norm: (u_x, u_y), product
// u_x^2 + u_y^2. Polynomial: (u), product sum_i c_i u^i.
template <int KIND>
struct Fields2D;
// NF fields leave a column kernel per instance; a row kernel transforms NIN
// fields back and forms NP products.
// Two-channel conservative convection: field u_c; inputs (u_0, u_1) of the
// trajectory; products u_c u_0, u_c u_1 (the terms of 1/2 sum_j i kappa_j
// F[u_c u_j]). Non-conservative: fields (u_c, d_x u_c, d_y u_c); inputs
// (u_0, u_1, d_x u_c, d_y u_c); product u_0 d_x u_c + u_1 d_y u_c. General:
// fields (u, u_x, u_y); products u^2 and u_x^2 + u_y^2. Coupled reactions
// (Gray-Scott, Belousov-Zhabotinsky): field u_c; inputs all channels of the
// trajectory; product the channel's reaction.
template <int KIND>
struct Fields2D {
  static constexpr bool COUPLED = KIND == NL_GRAY_SCOTT || KIND == NL_BZ;
  static constexpr int NF = KIND == NL_VORT ? 4 : KIND == NL_GRADNORM ? 2
                            : (KIND == NL_POLY || KIND == NL_CONV_CONS || COUPLED) ? 1 : 3;
  static constexpr int NIN = (KIND == NL_CONV_CONS || KIND == NL_GRAY_SCOTT) ? 2
                             : KIND == NL_BZ ? 3 : KIND == NL_CONV_NONCONS ? 4 : NF;
  static constexpr int NP = (KIND == NL_CONV_CONS || KIND == NL_GENERAL) ? 2 : 1;
  static_assert(KIND == NL_VORT || KIND == NL_GRADNORM || KIND == NL_POLY || KIND == NL_CONV_CONS ||
                    KIND == NL_CONV_NONCONS || KIND == NL_GENERAL || COUPLED,
                "unsupported 2-D nonlinear kind");
};

// ------------------------------------------------------------ row kernels
// R rows per workgroup; each row's real FFT of N points (M = N/2 complex) is
// owned by T = M/PT lanes of one wavefront.
template <int N>
struct RowCfg {
  static constexpr int M = N / 2;
  // 4 points per lane at small N: twice the wavefronts per launch and a
  // shorter per-lane chain for these latency-bound sizes
  static constexpr int PT = N <= EXMI_ROW_PT4_MAX ? 4 : 8;
  static constexpr int T = M / PT;
  static constexpr int R = EXMI_ROW_R;  // rows per workgroup
  static constexpr int WG = R * T;
  static constexpr int NY = M + 1;
  static constexpr int LD = R + 1;  // padded tile row
  static_assert(T <= 64, "row transform must fit one wavefront");
};

// Global accesses of the row/column interface buffers (T, F). COH = true
// (the cooperative rollout, where another workgroup of the same launch reads
// what this one wrote after a barrier) makes each access coherent across the
// device's L2 caches by itself: a relaxed agent-scope 64-bit atomic load or
// store, which the hardware serves at the device's point of coherence, so
// the barrier needs no cache-wide writeback or invalidate.
template <bool COH>
__device__ __forceinline__ cf gld(const cf* p) {
  if constexpr (COH) {
    const unsigned long long v =
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return cf{__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32))};
  } else {
    return *p;
  }
}

template <bool COH>
__device__ __forceinline__ void gst(cf* p, cf v) {
  if constexpr (COH) {
    const unsigned long long w =
        (unsigned long long)__float_as_uint(v.re) | ((unsigned long long)__float_as_uint(v.im) << 32);
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    *p = v;
  }
}

// Tile [ky][r] (padded) of R rows <-> column-major global field rows x0..x0+R.
template <int N, bool COH = false>
__device__ __forceinline__ void tile_load(cf* tl, const cf* __restrict__ g, int x0, int kc) {
  using C = RowCfg<N>;
  for (int i = threadIdx.x; i < C::NY * C::R; i += C::WG) {
    const int ky = i / C::R, r = i % C::R;
    tl[ky * C::LD + r] = ky < kc ? gld<COH>(g + (long long)ky * N + x0 + r) : cf{0.f, 0.f};
  }
}

template <int N, bool COH = false>
__device__ __forceinline__ void tile_store(const cf* tl, cf* __restrict__ g, int x0, int kc) {
  using C = RowCfg<N>;
  for (int i = threadIdx.x; i < kc * C::R; i += C::WG) {
    const int ky = i / C::R, r = i % C::R;
    gst<COH>(g + (long long)ky * N + x0 + r, tl[ky * C::LD + r]);
  }
}

// Register-staged copy of one [ky][x0, x0 + R) tile (ky < kc), for software
// pipelining: load() issues the global reads, store() writes them to the
// LDS tile later, so the next field's reads overlap the current transform.
template <int N, bool COH = false>
struct TileRegs {
  using C = RowCfg<N>;
  static constexpr int CNT = (C::NY * C::R + C::WG - 1) / C::WG;
  cf v[CNT];
  __device__ __forceinline__ void load(const cf* __restrict__ g, int x0, int kc) {
#pragma unroll
    for (int q = 0; q < CNT; ++q) {
      const int i = threadIdx.x + q * C::WG;
      const int ky = i / C::R, r = i % C::R;
      v[q] = (i < C::NY * C::R && ky < kc) ? gld<COH>(g + (long long)ky * N + x0 + r) : cf{0.f, 0.f};
    }
  }
  __device__ __forceinline__ void store(cf* tl) const {
#pragma unroll
    for (int q = 0; q < CNT; ++q) {
      const int i = threadIdx.x + q * C::WG;
      if (i < C::NY * C::R) tl[(i / C::R) * C::LD + i % C::R] = v[q];
    }
  }
};

template <int N>
__device__ __forceinline__ void tile_to_row(const cf* tl, int r, int j, cf (*z)[RowCfg<N>::PT], cf* nyq) {
  using C = RowCfg<N>;
#pragma unroll
  for (int s = 0; s < C::PT; ++s) z[0][s] = tl[(j + s * C::T) * C::LD + r];
  nyq[0] = tl[C::M * C::LD + r];
}

template <int N>
__device__ __forceinline__ void row_to_tile(cf* tl, int r, int j, const cf (*z)[RowCfg<N>::PT], const cf* nyq) {
  using C = RowCfg<N>;
#pragma unroll
  for (int s = 0; s < C::PT; ++s) tl[(j + s * C::T) * C::LD + r] = z[0][s];
  if (j == 0) tl[C::M * C::LD + r] = nyq[0];
}

template <int N>
struct RowLds {
  static constexpr int X = RowCfg<N>::R * LdsSize<RowCfg<N>::M>::value;  // transforms' exchange regions
  static constexpr int TL = RowCfg<N>::NY * RowCfg<N>::LD;               // transpose tile
};

// One row workgroup's pass: rows x0 = bx R .. x0 + R of instance inst.
template <int N, int MODE, int KIND, bool COH>
__device__ __forceinline__ void row2d_body(const Step2DArgs& a, int bx, int inst, cf* lds_x, cf* lds_t) {
  using C = RowCfg<N>;
  constexpr int M = C::M, PT = C::PT, T = C::T, R = C::R;
  constexpr int LDSN = LdsSize<M>::value;

  const int x0 = bx * R;
  const int r = threadIdx.x / T, j = threadIdx.x % T;
  const int x = x0 + r;
  cf* lds = lds_x + r * LDSN;
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> tr;
  tr.load(j, RealTables{a.tw_row, a.fw_row});
  const long long fstride = a.tstride;  // T of one instance
  const int kc = a.kmax + 1 < C::NY ? a.kmax + 1 : C::NY;
  // physical row x of this instance in a (batch, C, N, N) state
  const long long urow = (long long)(inst % a.channels) * N * N + (long long)x * N;
  const int traj = inst / a.channels;
  constexpr float NN = (float)N;
  constexpr float rn2 = 1.f / (2.f * NN * NN);  // inverse of 2 u_hat gives 2 N^2 u

  cf z[1][PT];
  cf nyq[1];
  if constexpr (MODE == ROW_RA) {
    const float2* u = reinterpret_cast<const float2*>(a.u_in + (long long)traj * a.u_in_bstride + urow);
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const float2 v = u[j + s * T];
      z[0][s] = cf{v.x, v.y};
    }
    rfft<M, PT, 1>(z, nyq, ln, lds, tr);
    row_to_tile<N>(lds_t, r, j, z, nyq);
    __syncthreads();
    tile_store<N, COH>(lds_t, a.T + (long long)inst * fstride, x0, C::NY);
  }
  using FD = Fields2D<KIND>;
  constexpr int NF = FD::NF, NIN = FD::NIN, NP = FD::NP;
  const long long fsz = (long long)kc * N;
  const int ch = inst % a.channels;
  // input f of the products: this instance's field f, or for the
  // convections the velocity field u_j of channel j of the same trajectory
  auto src = [&](int f) -> const cf* {
    int i = inst, fld = f;
    if constexpr (KIND == NL_CONV_CONS || FD::COUPLED) {
      i = inst - ch + f;
      fld = 0;
    } else if constexpr (KIND == NL_CONV_NONCONS) {
      if (f < 2) {
        i = inst - ch + f;
        fld = 0;
      } else {
        fld = f - 1;
      }
    }
    return a.F + ((long long)i * NF + fld) * fsz;
  };
  TileRegs<N, COH> pf;
  // RCB: the first field's reads are issued before RC, whose transform they overlap
  if constexpr (MODE == ROW_RCB) pf.load(src(0), x0, kc);
  if constexpr (MODE == ROW_RC || MODE == ROW_RCB) {
    tile_load<N, COH>(lds_t, a.T + (long long)inst * fstride, x0, C::NY);
    __syncthreads();
    tile_to_row<N>(lds_t, r, j, z, nyq);
    irfft<M, PT, 1>(z, nyq, ln, lds, tr);
#pragma unroll
    for (int s = 0; s < PT; ++s) z[0][s] = cf{z[0][s].re * rn2, z[0][s].im * rn2};
    float2* u = reinterpret_cast<float2*>(a.u_out + (long long)traj * a.u_out_bstride + urow);
#pragma unroll
    for (int s = 0; s < PT; ++s) u[j + s * T] = make_float2(z[0][s].re, z[0][s].im);
    if (MODE == ROW_RC && a.write_T) {
      // the next step's forward transform of the float32 state just written
      rfft<M, PT, 1>(z, nyq, ln, lds, tr);
      __syncthreads();
      row_to_tile<N>(lds_t, r, j, z, nyq);
      __syncthreads();
      tile_store<N, COH>(lds_t, a.T + (long long)inst * fstride, x0, C::NY);
    }
    // RCB: the products below overwrite T's rows ky < kc of the same x
    // range, which every thread of the workgroup has read into LDS by now
    if constexpr (MODE == ROW_RCB) __syncthreads();
  }
  if constexpr (MODE == ROW_RB || MODE == ROW_RCB) {
    cf A[PT], B2[PT], P[NP][PT];
    if constexpr (MODE == ROW_RB) pf.load(src(0), x0, kc);
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      if (f > 0) __syncthreads();
      pf.store(lds_t);
      __syncthreads();
      if (f + 1 < NIN) pf.load(src(f + 1), x0, kc);  // overlaps this field's transform
      tile_to_row<N>(lds_t, r, j, z, nyq);
      irfft<M, PT, 1>(z, nyq, ln, lds, tr);
      if constexpr (KIND == NL_GRADNORM) {
        if (a.Gout) {  // the adjoint's gradient fields (as its RBG/RG passes store them)
          float2* G = reinterpret_cast<float2*>(a.Gout + ((long long)inst * 4 + a.gstage * 2 + f) * N * N +
                                                (long long)x * N);
#pragma unroll
          for (int s = 0; s < PT; ++s) G[j + s * T] = make_float2(z[0][s].re, z[0][s].im);
        }
      }
      if constexpr (FD::COUPLED) {
        // the channel's reaction at unit scale (inputs are 2 N^2 u)
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          if (f == 0) {
            A[s] = q;
          } else if (f + 1 < NIN) {
            B2[s] = q;
          } else {
            float r[3][2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const float u0 = (h ? A[s].im : A[s].re) * rn2;
              if constexpr (KIND == NL_GRAY_SCOTT) {
                gray_scott(u0, (h ? q.im : q.re) * rn2, a.feed, a.kill, r[0][h], r[1][h]);
                r[2][h] = 0.f;
              } else {
                belousov_zhabotinsky(u0, (h ? B2[s].im : B2[s].re) * rn2, (h ? q.im : q.re) * rn2, r[0][h], r[1][h],
                                     r[2][h]);
              }
            }
            P[0][s] = ch == 0 ? cf{r[0][0], r[0][1]} : ch == 1 ? cf{r[1][0], r[1][1]} : cf{r[2][0], r[2][1]};
          }
        }
      } else if constexpr (KIND == NL_CONV_CONS) {
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          if (f == 0) {
            A[s] = q;
          } else {
            const cf uc = ch == 0 ? A[s] : q;
            P[0][s] = cf{uc.re * A[s].re, uc.im * A[s].im};
            P[1][s] = cf{uc.re * q.re, uc.im * q.im};
          }
        }
      } else if constexpr (KIND == NL_CONV_NONCONS) {
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          if (f == 0) {
            A[s] = q;
          } else if (f == 1) {
            B2[s] = q;
          } else if (f == 2) {
            P[0][s] = cf{A[s].re * q.re, A[s].im * q.im};
          } else {
            P[0][s] = cf{fmaf(B2[s].re, q.re, P[0][s].re), fmaf(B2[s].im, q.im, P[0][s].im)};
          }
        }
      } else if constexpr (KIND == NL_GENERAL) {
        // u^2 and u_x^2 + u_y^2; with prod_nc the first product is
        // u (u_x + u_y) (nonlin_fun/_convection.py:204-231)
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          if (f == 0) {
            A[s] = q;
            P[0][s] = cf{q.re * q.re, q.im * q.im};
          } else if (f == 1) {
            P[1][s] = cf{q.re * q.re, q.im * q.im};
            B2[s] = cf{A[s].re * q.re, A[s].im * q.im};
          } else {
            P[1][s] = cf{fmaf(q.re, q.re, P[1][s].re), fmaf(q.im, q.im, P[1][s].im)};
            const cf nc = cf{fmaf(A[s].re, q.re, B2[s].re), fmaf(A[s].im, q.im, B2[s].im)};
            if (a.prod_nc) P[0][s] = nc;
          }
        }
      } else if constexpr (KIND == NL_VORT) {
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          if (f == 0 || f == 2) {
            A[s] = q;
          } else if (f == 1) {
            P[0][s] = cf{A[s].re * q.re, A[s].im * q.im};
          } else {
            P[0][s] = cf{fmaf(A[s].re, q.re, P[0][s].re), fmaf(A[s].im, q.im, P[0][s].im)};
          }
        }
      } else if constexpr (KIND == NL_POLY) {
        // the same power loop as the reference, on u = (2 N^2 u) / (2 N^2)
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const float ur = z[0][s].re * rn2, ui = z[0][s].im * rn2;
          float pr = 1.f, pi = 1.f, ar = 0.f, ai = 0.f;
          for (int i = 0; i < a.npoly; ++i) {
            ar += a.poly[i] * pr;
            ai += a.poly[i] * pi;
            pr *= ur;
            pi *= ui;
          }
          P[0][s] = cf{ar, ai};
        }
      } else {
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          P[0][s] = f == 0 ? cf{q.re * q.re, q.im * q.im}
                           : cf{fmaf(q.re, q.re, P[0][s].re), fmaf(q.im, q.im, P[0][s].im)};
        }
      }
    }
    (void)A;
    (void)B2;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int s = 0; s < PT; ++s) z[0][s] = P[p][s];
      rfft<M, PT, 1>(z, nyq, ln, lds, tr, a.kmax);  // only ky <= kmax are kept
      __syncthreads();
      row_to_tile<N>(lds_t, r, j, z, nyq);
      __syncthreads();
      tile_store<N, COH>(lds_t, a.T + (long long)inst * fstride + p * fsz, x0, kc);
    }
  }
}

template <int N, int MODE, int KIND>
__global__ void __launch_bounds__(RowCfg<N>::WG) row2d_kernel(Step2DArgs a) {
  __shared__ cf lds_x[RowLds<N>::X];
  __shared__ cf lds_t[RowLds<N>::TL];
  row2d_body<N, MODE, KIND, false>(a, blockIdx.x, blockIdx.y, lds_x, lds_t);
}

template <int N, int MODE, int KIND>
inline hipError_t launch_row2d(const Step2DArgs& a, int ninst, hipStream_t st) {
  using C = RowCfg<N>;
  hipLaunchKernelGGL((row2d_kernel<N, MODE, KIND>), dim3(N / C::R, ninst), dim3(C::WG), 0, st, a);
  return hipGetLastError();
}

// --------------------------------------------------------- column kernels
// Each column (fixed ky, all N x-values, contiguous) is a complex N-point FFT
// owned by T = N/PT lanes; CPB columns per 256-thread workgroup.
template <int N>
struct ColCfg {
  static constexpr int PT = N <= EXMI_COL_PT4_MAX ? 4 : 8;
  static constexpr int T = N / PT;
  static constexpr int WG = 256;
  static constexpr int CPB = WG / T;  // columns per workgroup
  static constexpr int NY = N / 2 + 1;
  static_assert(T <= 64, "column transform must fit one wavefront");
};

template <int N>
struct ColLds {
  static constexpr int X = ColCfg<N>::CPB * LdsSize<N>::value;
};

// One column workgroup's pass: columns ky = bx CPB .. of instance inst.
template <int N, int MODE, int KIND, bool COH>
__device__ __forceinline__ void col2d_body(const Step2DArgs& a, int bx, int inst, cf* lds_x) {
  using C = ColCfg<N>;
  constexpr int PT = C::PT, T = C::T;
  constexpr int LDSN = LdsSize<N>::value;
  constexpr int NF = Fields2D<KIND>::NF;
  constexpr float NN = (float)N;
  // 2 N_hat from the forward transform of the product (8 N^4 F[product])
  constexpr float q2 = 2.f / (8.f * NN * NN * NN * NN);

  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const int ky_raw = bx * C::CPB + c;
  const bool active = ky_raw < C::NY;
  const int ky = active ? ky_raw : 0;  // inactive groups mirror column 0, store nothing
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, a.tw_col);

  const int kc = a.kmax + 1 < C::NY ? a.kmax + 1 : C::NY;
  const bool kept_col = ky < kc;
  const long long full = (long long)C::NY * N;
  const float kap_y = a.kscale * (float)ky;

  // per-slot signed kx and the dealias predicate |kx| <= kmax
  float kap_x[PT];
  bool keep_x[PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int idx = j + s * T;
    const int kx = idx < N / 2 ? idx : idx - N;
    kap_x[s] = a.kscale * (float)kx;
    keep_x[s] = (kx < 0 ? -kx : kx) <= a.kmax;
  }

  // forcing (reference scaling) carried as 2 N_hat
  auto forcing = [&](int s) -> cf {
    return (a.inj_mode == ky && j + s * T == 0) ? cf{2.f * a.inj_value, 0.f} : cf{0.f, 0.f};
  };

  // nonlinear term 2 N_hat of this column from the products (ky < kc)
  auto nonlinear = [&](cf* Nh) {
    if (kept_col) {
      constexpr int NP = Fields2D<KIND>::NP;
      cf v[NP][PT];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const cf* pp = a.T + (long long)inst * a.tstride + ((long long)p * kc + ky) * N;
#pragma unroll
        for (int s = 0; s < PT; ++s) v[p][s] = gld<COH>(pp + j + s * T);
      }
      cfft<N, PT, false, NP>(v, ln, lds, tr);
      // the polynomial product is formed at unit scale, so its forward
      // transforms give 2 N_hat directly (no sign: nonlin_fun/_polynomial.py:76)
      const float cs = (KIND == NL_VORT || KIND == NL_CONV_NONCONS) ? -a.scale * q2
                       : (KIND == NL_POLY || Fields2D<KIND>::COUPLED) ? 1.f
                                                                      : -a.scale * 0.5f * q2;
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        cf r;
        if constexpr (KIND == NL_CONV_CONS) {
          // -b/2 (i kx F[u_c u_0] + i ky F[u_c u_1])
          const cf w0 = v[0][s], w1 = v[NP - 1][s];
          r = cs * cf{-(kap_x[s] * w0.im + kap_y * w1.im), fmaf(kap_x[s], w0.re, kap_y * w1.re)};
        } else if constexpr (KIND == NL_GENERAL) {
          // g0 F[u^2] + (g1/2) i (kx + ky) F[u^2] + (g2/2) F[|grad u|^2] (mean-free)
          const cf w0 = v[0][s], w1 = v[NP - 1][s];
          const float ck = 0.5f * a.g1 * q2 * (kap_x[s] + kap_y);
          const float cg = (a.zero_mode_fix && ky == 0 && j + s * T == 0) ? 0.f : 0.5f * a.g2 * q2;
          const float c0 = a.g0 * q2;
          r = cf{fmaf(c0, w0.re, fmaf(-ck, w0.im, cg * w1.re)), fmaf(c0, w0.im, fmaf(ck, w0.re, cg * w1.im))};
        } else if constexpr (KIND == NL_POLY) {
          // scale * Laplace F[u^3] = -scale (kx^2 + ky^2) F[u^3] for Cahn-Hilliard
          const float lap = -(kap_x[s] * kap_x[s] + kap_y * kap_y);
          r = a.poly_lap ? (a.scale * lap) * v[0][s] : v[0][s];
        } else {
          r = cs * v[0][s];
          if (KIND == NL_GRADNORM && a.zero_mode_fix && ky == 0 && j + s * T == 0) r = cf{0.f, 0.f};
        }
        Nh[s] = keep_x[s] ? r : cf{0.f, 0.f};
        Nh[s] = Nh[s] + forcing(s);
      }
    } else {
#pragma unroll
      for (int s = 0; s < PT; ++s) Nh[s] = forcing(s);
    }
  };

  // fields of the nonlinear evaluation of state Sx (2 u_hat), inverse
  // transformed along x and stored (ky < kc only)
  auto fields_out = [&](const cf* Sx) {
    if (!kept_col || !active) return;
    cf v[NF][PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf w = keep_x[s] ? Sx[s] : cf{0.f, 0.f};
      const float kx = kap_x[s], kyv = kap_y;
      if constexpr (KIND == NL_VORT) {
        // psi = Laplace^-1 omega (1 at the mean mode), u = i ky psi,
        // v = -i kx psi, omega_x = i kx omega, omega_y = i ky omega
        const float lap = -(kx * kx + kyv * kyv);
        const float inv = lap == 0.f ? 1.f : 1.f / lap;
        const cf psi = inv * w;
        v[0][s] = cf{-kyv * psi.im, kyv * psi.re};
        v[1][s] = cf{-kx * w.im, kx * w.re};
        v[2][s] = cf{kx * psi.im, -kx * psi.re};
        v[3][s] = cf{-kyv * w.im, kyv * w.re};
      } else if constexpr (KIND == NL_POLY || KIND == NL_CONV_CONS || Fields2D<KIND>::COUPLED) {
        v[0][s] = w;
      } else if constexpr (KIND == NL_CONV_NONCONS || KIND == NL_GENERAL) {
        v[0][s] = w;
        v[1][s] = cf{-kx * w.im, kx * w.re};
        v[2][s] = cf{-kyv * w.im, kyv * w.re};
      } else {
        v[0][s] = cf{-kx * w.im, kx * w.re};
        v[1][s] = cf{-kyv * w.im, kyv * w.re};
      }
    }
    cfft<N, PT, true, NF>(v, ln, lds, tr);
    cf* F = a.F + (long long)inst * NF * kc * N;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      cf* o = F + ((long long)f * kc + ky) * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) gst<COH>(o + j + s * T, v[f][s]);
    }
  };

  const int traj = inst / a.channels, chan = inst % a.channels;
  const float* tb =
      a.tables + ((long long)(a.lin_batch > 1 ? traj : 0) * a.lin_channels + (a.lin_channels > 1 ? chan : 0)) *
                     a.ncoef * full;
  auto coef = [&](int i, int s) { return tb[(long long)i * full + (long long)ky * N + j + s * T]; };
  // E x with E = exp(L dt), complex when L is (odd derivatives)
  // coefficient i (0: E, 1: E_half for orders 3-4) times x, complex when
  // the plan's L is
  auto cmul = [&](int i, int s, cf x) -> cf {
    const float er = coef(i, s);
    if (a.e_im < 0) return er * x;
    const float ei = coef(a.e_im + i, s);
    return cf{fmaf(er, x.re, -ei * x.im), fmaf(er, x.im, ei * x.re)};
  };
  auto emul = [&](int s, cf x) -> cf { return cmul(0, s, x); };
  // the last stage's new state 2 u_hat' of this column: inverse column
  // transform into T (physical steps; RC follows), or u_hat' itself into
  // spec_out [kx][ky] (Fourier-space steps)
  auto finish = [&](cf (*v)[PT]) {
    if (a.spec_io) {
      if (active) {
        cf* o = a.spec_out + (long long)inst * a.spec_bstride + ky;
#pragma unroll
        for (int s = 0; s < PT; ++s) o[(long long)(j + s * T) * C::NY] = 0.5f * v[0][s];
      }
      return;
    }
    cfft<N, PT, true, 1>(v, ln, lds, tr);
    if (active) {
      cf* t = a.T + (long long)inst * a.tstride + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) gst<COH>(t + j + s * T, v[0][s]);
    }
  };

  if constexpr (MODE == COL_CA) {
    cf v[1][PT];
    const cf* t = a.T + (long long)inst * a.tstride + (long long)ky * N;
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = gld<COH>(t + j + s * T);
    cfft<N, PT, false, 1>(v, ln, lds, tr);
    if (active) {
      cf* s0 = a.S0 + (long long)inst * full + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) s0[j + s * T] = v[0][s];
    }
    fields_out(v[0]);
  } else if constexpr (MODE == COL_CB) {
    cf Nh[PT], S[PT];
    // the state and the coefficients are read before the products, so
    // their latency overlaps the products' transform
    const cf* s0 = a.S0 + (long long)inst * full + (long long)ky * N;
    float c1[PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      S[s] = emul(s, s0[j + s * T]);
      c1[s] = coef(1, s);
    }
    nonlinear(Nh);
#pragma unroll
    for (int s = 0; s < PT; ++s) S[s] = S[s] + c1[s] * Nh[s];
    if (active) {
      cf* s1 = a.S1 + (long long)inst * full + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) s1[j + s * T] = S[s];
      if (kept_col) {
        cf* n0 = a.N0 + ((long long)inst * kc + ky) * N;
#pragma unroll
        for (int s = 0; s < PT; ++s) n0[j + s * T] = Nh[s];
      }
    }
    fields_out(S);
  } else if constexpr (MODE >= COL_C3A && MODE <= COL_C4D) {
    // column entries of a full [ky][kx] buffer, and N0 (forcing outside the
    // kept columns)
    auto ld = [&](const cf* buf, cf* v) {
      const cf* p = buf + (long long)inst * full + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) v[s] = p[j + s * T];
    };
    auto st = [&](cf* buf, const cf* v) {
      if (!active) return;
      cf* p = buf + (long long)inst * full + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) p[j + s * T] = v[s];
    };
    auto ld_n0 = [&](cf* v) {
      const cf* n0 = a.N0 + ((long long)inst * kc + ky) * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) v[s] = kept_col ? n0[j + s * T] : forcing(s);
    };
    cf Nh[PT], U0[PT], Ac[PT], S[PT];
    nonlinear(Nh);
    ld(a.S0, U0);
    if constexpr (MODE == COL_C3A || MODE == COL_C4A) {
      // A = E U0 + (c3 | c4) N0; S = E_half U0 + c1 N0
      constexpr int cA = MODE == COL_C3A ? 4 : 3;
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        Ac[s] = cmul(0, s, U0[s]) + coef(cA, s) * Nh[s];
        S[s] = cmul(1, s, U0[s]) + coef(2, s) * Nh[s];
      }
      st(a.A, Ac);
      if (active && kept_col) {
        cf* n0 = a.N0 + ((long long)inst * kc + ky) * N;
#pragma unroll
        for (int s = 0; s < PT; ++s) n0[j + s * T] = Nh[s];
      }
      if constexpr (MODE == COL_C4A) st(a.S1, S);
      fields_out(S);
    } else if constexpr (MODE == COL_C3B) {
      // A += c4 N1; S = E U0 + c2 (2 N1 - N0)
      cf N0v[PT];
      ld(a.A, Ac);
      ld_n0(N0v);
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        Ac[s] = Ac[s] + coef(5, s) * Nh[s];
        S[s] = cmul(0, s, U0[s]) + coef(3, s) * (2.f * Nh[s] - N0v[s]);
      }
      st(a.A, Ac);
      fields_out(S);
    } else if constexpr (MODE == COL_C4B || MODE == COL_C4C) {
      // A += 2 c5 N_k; stage 2: S = E_half U0 + c1 N1;
      // stage 3: S = E_half S1 + c1 (2 N2 - N0)
      ld(a.A, Ac);
#pragma unroll
      for (int s = 0; s < PT; ++s) Ac[s] = Ac[s] + (coef(4, s) * 2.f) * Nh[s];
      st(a.A, Ac);
      if constexpr (MODE == COL_C4B) {
#pragma unroll
        for (int s = 0; s < PT; ++s) S[s] = cmul(1, s, U0[s]) + coef(2, s) * Nh[s];
      } else {
        cf N0v[PT], S1v[PT];
        ld_n0(N0v);
        ld(a.S1, S1v);
#pragma unroll
        for (int s = 0; s < PT; ++s) S[s] = cmul(1, s, S1v[s]) + coef(2, s) * (2.f * Nh[s] - N0v[s]);
      }
      fields_out(S);
    } else {
      // last stage: S' = A + (c5 | c6) N_last -> inverse column transform -> T
      constexpr int cL = MODE == COL_C3C ? 6 : 5;
      cf v[1][PT];
      ld(a.A, Ac);
#pragma unroll
      for (int s = 0; s < PT; ++s) v[0][s] = Ac[s] + coef(cL, s) * Nh[s];
      finish(v);
    }
    (void)U0;
  } else if constexpr (MODE == COL_C0 || MODE == COL_C1) {
    cf v[1][PT];
    if constexpr (MODE == COL_C0) {
      if (a.spec_io) {
        const cf* in = a.spec + (long long)inst * a.spec_bstride + ky;
#pragma unroll
        for (int s = 0; s < PT; ++s) v[0][s] = 2.f * in[(long long)(j + s * T) * C::NY];
      } else {
        const cf* t = a.T + (long long)inst * a.tstride + (long long)ky * N;
#pragma unroll
        for (int s = 0; s < PT; ++s) v[0][s] = gld<COH>(t + j + s * T);
        cfft<N, PT, false, 1>(v, ln, lds, tr);
      }
#pragma unroll
      for (int s = 0; s < PT; ++s) v[0][s] = emul(s, v[0][s]);
    } else {
      cf Nh[PT];
      nonlinear(Nh);
      const cf* s0 = a.S0 + (long long)inst * full + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) v[0][s] = emul(s, s0[j + s * T]) + coef(1, s) * Nh[s];
    }
    finish(v);
  } else if constexpr (MODE == COL_CA_SPEC) {
    // CA from a given spectrum u_hat [kx][ky] (exmi_nonlinear): S0 = 2 u_hat
    cf v[1][PT];
    const cf* in = a.spec + (long long)inst * a.spec_bstride + ky;
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = 2.f * in[(long long)(j + s * T) * C::NY];
    if (active) {
      cf* s0 = a.S0 + (long long)inst * full + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) s0[j + s * T] = v[0][s];
    }
    fields_out(v[0]);
  } else if constexpr (MODE == COL_NOUT) {
    // the nonlinear term N_hat of the products in T, written as [kx][ky]
    cf Nh[PT];
    nonlinear(Nh);
    if (active) {
      cf* o = a.spec_out + (long long)inst * a.spec_bstride + ky;
#pragma unroll
      for (int s = 0; s < PT; ++s) o[(long long)(j + s * T) * C::NY] = 0.5f * Nh[s];
    }
  } else if constexpr (MODE == COL_FWD) {
    // rfftn: the row spectra of column ky -> X[kx][ky] (rows gave 2 X)
    cf v[1][PT];
    const cf* t = a.T + (long long)inst * a.tstride + (long long)ky * N;
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = gld<COH>(t + j + s * T);
    cfft<N, PT, false, 1>(v, ln, lds, tr);
    if (active) {
      cf* o = a.spec + (long long)inst * a.spec_bstride + ky;
#pragma unroll
      for (int s = 0; s < PT; ++s) o[(long long)(j + s * T) * C::NY] = 0.5f * v[0][s];
    }
  } else if constexpr (MODE == COL_INV) {
    // irfftn: X[kx][ky] -> inverse column transform of 2 X -> T (RC then
    // finishes with the row c2r and the 1/(2 N^2) scaling)
    cf v[1][PT];
    const cf* in = a.spec + (long long)inst * a.spec_bstride + ky;
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = 2.f * in[(long long)(j + s * T) * C::NY];
    cfft<N, PT, true, 1>(v, ln, lds, tr);
    if (active) {
      cf* t = a.T + (long long)inst * a.tstride + (long long)ky * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) gst<COH>(t + j + s * T, v[0][s]);
    }
  } else {
    // CC: S' = S1 + c2 (N1 - N0) -> inverse column transform -> T. CCA
    // also keeps S' as the next step's S0 and, with next_fields, forms
    // its fields (the next step's CA)
    cf Nh[PT], v[1][PT], N0s[PT];
    float c2[PT];
    // S1, N0 and the coefficients are read before the products (as in CB)
    const cf* s1 = a.S1 + (long long)inst * full + (long long)ky * N;
    const cf* n0 = a.N0 + ((long long)inst * kc + ky) * N;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      N0s[s] = kept_col ? n0[j + s * T] : forcing(s);
      v[0][s] = s1[j + s * T];
      c2[s] = coef(2, s);
    }
    nonlinear(Nh);
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = v[0][s] + c2[s] * (Nh[s] - N0s[s]);
    if constexpr (MODE == COL_CCA) {
      if (active) {
        cf* s0 = a.S0 + (long long)inst * full + (long long)ky * N;
#pragma unroll
        for (int s = 0; s < PT; ++s) s0[j + s * T] = v[0][s];
      }
      if (a.next_fields) fields_out(v[0]);
    }
    finish(v);
  }
}

template <int N, int MODE, int KIND>
__global__ void __launch_bounds__(ColCfg<N>::WG) col2d_kernel(Step2DArgs a) {
  __shared__ cf lds_x[ColLds<N>::X];
  col2d_body<N, MODE, KIND, false>(a, blockIdx.x, blockIdx.y, lds_x);
}

template <int N, int MODE, int KIND>
inline hipError_t launch_col2d(const Step2DArgs& a, int ninst, hipStream_t st) {
  using C = ColCfg<N>;
  const int groups = (C::NY + C::CPB - 1) / C::CPB;
  hipLaunchKernelGGL((col2d_kernel<N, MODE, KIND>), dim3(groups, ninst), dim3(C::WG), 0, st, a);
  return hipGetLastError();
}

}  // namespace exmi
