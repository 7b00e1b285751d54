// Adjoint (vector-Jacobian product) of the 2-D ETDRK2 step with the
// gradient-norm nonlinearity (KuramotoSivashinsky in 2-D). The reference
// differentiates its step with jax.grad through lax.scan (_utils.py:251);
// here the transpose is written out.
//
// The inner products are sum(f g) for fields and the Parseval form for
// spectra. Under them rfftn and irfftn are mutual adjoints, a real diagonal
// factor (E, c1, c2, the dealias mask) is its own adjoint, i kappa has adjoint
// -i kappa, and removing the mean is self-adjoint. With N(u) =
// -b/2 P Z F[|grad u|^2] (P = dealias, Z = mean removal, _gradient_norm.py:84-101),
// g_d = F^-1[P i kappa_d u_hat] and h = F^-1[P Z mu]:
//
//   J(u)^T mu = -b sum_d P (-i kappa_d) F[g_d h]
//
// and for ETDRK2 (_etdrk_2.py:91-102), with cotangent lam of the new state:
//
//   rho = lam + J(S1)^T (c2 lam),  lam_in = E rho + J(u_hat)^T (c1 rho - c2 lam).
//
// Launch chain of one transposed step (B instances; the forward kernels of
// step2d.hpp plus the modes below):
//   recompute : RA(u) CA RBG(stage 0) CB RG(stage 1)  -> gradient fields G
//   transpose : RA(g) CL RP(stage 1) CD RP(stage 0) CE RC -> g_in
//
// Scalings follow step2d.hpp: spectra are carried as 2x, physical fields from
// the inverse chain as 2 N^2 x, so the products come back from the forward
// chain as 8 N^4 F[.] and the factor q2 = 2 / (8 N^4) restores 2x (exact).
#pragma once
#include "step2d.hpp"

namespace exmi {

struct Adj2DArgs {
  Step2DArgs s;
  float* G;   // (ninst, 2 stages, 2 fields, N, N) physical gradient fields (scaled 2 N^2)
  int stage;  // 0: fields of u_hat, 1: fields of the stage state S1
  cf* LAM;    // (ninst, NY, N) cotangent of the new state's spectrum [ky][kx] (2 lam)
  cf* RHO;    // (ninst, NY, N)
  cf* H;      // (ninst, Kc, N) masked cotangent after the inverse transform along x, [ky][x]
  cf* P;      // (ninst, 2, Kc, N) product spectra along y, [ky][x]
};

enum Row2DAdjMode { ROWA_RBG = 0, ROWA_RG = 1, ROWA_RP = 2 };
enum Col2DAdjMode { COLA_CL = 0, COLA_CD = 1, COLA_CE = 2 };

// RBG: the forward RB (gradient norm) that also keeps the physical gradient
// fields; RG: the fields only; RP: the products g_d h transformed along y.
// One row workgroup's pass (rows x0 = bx R .. of instance inst); COH as in
// step2d.hpp (the buffers other workgroups read or wrote: F, T, H, P).
template <int N, int MODE, bool COH>
__device__ __forceinline__ void row2d_adj_body(const Adj2DArgs& aa, int bx, int inst, cf* lds_x, cf* lds_t) {
  using C = RowCfg<N>;
  constexpr int M = C::M, PT = C::PT, T = C::T, R = C::R;
  constexpr int LDSN = LdsSize<M>::value;
  const Step2DArgs& a = aa.s;

  const int x0 = bx * R;
  const int r = threadIdx.x / T, j = threadIdx.x % T;
  const int x = x0 + r;
  cf* lds = lds_x + r * LDSN;
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> tr;
  tr.load(j, RealTables{a.tw_row, a.fw_row});
  const long long fstride = (long long)C::NY * N;
  const int kc = a.kmax + 1 < C::NY ? a.kmax + 1 : C::NY;
  constexpr long long NN2 = (long long)N * N / 2;  // one field in float2
  float2* G = reinterpret_cast<float2*>(aa.G + ((long long)inst * 4 + aa.stage * 2) * N * N + (long long)x * N);

  cf z[1][PT];
  cf nyq[1];
  if constexpr (MODE == ROWA_RBG || MODE == ROWA_RG) {
    const cf* F = a.F + (long long)inst * 2 * kc * N;
    cf P[PT];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      if (f > 0) __syncthreads();
      tile_load<N, COH>(lds_t, F + (long long)f * kc * N, x0, kc);
      __syncthreads();
      tile_to_row<N>(lds_t, r, j, z, nyq);
      irfft<M, PT, 1>(z, nyq, ln, lds, tr);
#pragma unroll
      for (int s = 0; s < PT; ++s) G[f * NN2 + j + s * T] = make_float2(z[0][s].re, z[0][s].im);
      if constexpr (MODE == ROWA_RBG) {
        // same arithmetic as the forward RB
#pragma unroll
        for (int s = 0; s < PT; ++s) {
          const cf q = z[0][s];
          P[s] = f == 0 ? cf{q.re * q.re, q.im * q.im} : cf{fmaf(q.re, q.re, P[s].re), fmaf(q.im, q.im, P[s].im)};
        }
      }
    }
    if constexpr (MODE == ROWA_RBG) {
#pragma unroll
      for (int s = 0; s < PT; ++s) z[0][s] = P[s];
      rfft<M, PT, 1>(z, nyq, ln, lds, tr, a.kmax);
      __syncthreads();
      row_to_tile<N>(lds_t, r, j, z, nyq);
      __syncthreads();
      tile_store<N, COH>(lds_t, a.T + (long long)inst * fstride, x0, kc);
    }
    (void)P;
  } else {
    static_assert(MODE == ROWA_RP, "unknown adjoint row mode");
    tile_load<N, COH>(lds_t, aa.H + (long long)inst * kc * N, x0, kc);
    __syncthreads();
    tile_to_row<N>(lds_t, r, j, z, nyq);
    irfft<M, PT, 1>(z, nyq, ln, lds, tr);
    cf Hr[PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) Hr[s] = z[0][s];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        const float2 g = G[d * NN2 + j + s * T];
        z[0][s] = cf{g.x * Hr[s].re, g.y * Hr[s].im};
      }
      rfft<M, PT, 1>(z, nyq, ln, lds, tr, a.kmax);
      __syncthreads();
      row_to_tile<N>(lds_t, r, j, z, nyq);
      __syncthreads();
      tile_store<N, COH>(lds_t, aa.P + ((long long)inst * 2 + d) * kc * N, x0, kc);
    }
  }
}

template <int N, int MODE>
__global__ void __launch_bounds__(RowCfg<N>::WG) row2d_adj_kernel(Adj2DArgs aa) {
  __shared__ cf lds_x[RowLds<N>::X];
  __shared__ cf lds_t[RowLds<N>::TL];
  row2d_adj_body<N, MODE, false>(aa, blockIdx.x, blockIdx.y, lds_x, lds_t);
}

template <int N, int MODE>
inline hipError_t launch_row2d_adj(const Adj2DArgs& a, int ninst, hipStream_t st) {
  using C = RowCfg<N>;
  hipLaunchKernelGGL((row2d_adj_kernel<N, MODE>), dim3(N / C::R, ninst), dim3(C::WG), 0, st, a);
  return hipGetLastError();
}

// CL: lam = column transform of RA(g); H <- columns of P Z (c2 lam).
// CD: rho = lam + J(S1)^T (c2 lam); H <- columns of P Z (c1 rho - c2 lam).
// CE: lam_in = E rho + J(u_hat)^T (...); inverse transform along x -> T.
template <int N, int MODE, bool COH>
__device__ __forceinline__ void col2d_adj_body(const Adj2DArgs& aa, int bx, int inst, cf* lds_x) {
  using C = ColCfg<N>;
  constexpr int PT = C::PT, T = C::T;
  constexpr int LDSN = LdsSize<N>::value;
  constexpr float NN = (float)N;
  constexpr float q2 = 2.f / (8.f * NN * NN * NN * NN);
  const Step2DArgs& a = aa.s;

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
  const long long at = (long long)inst * full + (long long)ky * N;
  const float kap_y = a.kscale * (float)ky;
  const bool dc_col = a.zero_mode_fix && ky == 0;

  float kap_x[PT];
  bool keep_x[PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int idx = j + s * T;
    const int kx = idx < N / 2 ? idx : idx - N;
    kap_x[s] = a.kscale * (float)kx;
    keep_x[s] = (kx < 0 ? -kx : kx) <= a.kmax;
  }
  const float* tb = a.tables + (long long)(a.lin_batch > 1 ? inst : 0) * 3 * full;
  auto coef = [&](int i, int s) { return tb[(long long)i * full + (long long)ky * N + j + s * T]; };

  // H column: dealiased, mean-free cotangent X, inverse transformed along x
  auto h_out = [&](const cf* X) {
    if (!kept_col || !active) return;
    cf v[1][PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = (keep_x[s] && !(dc_col && j + s * T == 0)) ? X[s] : cf{0.f, 0.f};
    cfft<N, PT, true, 1>(v, ln, lds, tr);
    cf* h = aa.H + ((long long)inst * kc + ky) * N;
#pragma unroll
    for (int s = 0; s < PT; ++s) gst<COH>(h + j + s * T, v[0][s]);
  };

  // 2 J^T mu of this column from the product spectra (zero outside the kept
  // columns): -b P sum_d (-i kappa_d) F[g_d h]
  auto adj_nonlinear = [&](cf* A) {
    if (kept_col) {
      cf v[2][PT];
      const cf* p0 = aa.P + ((long long)inst * 2 * kc + ky) * N;
      const cf* p1 = p0 + (long long)kc * N;
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        v[0][s] = gld<COH>(p0 + j + s * T);
        v[1][s] = gld<COH>(p1 + j + s * T);
      }
      cfft<N, PT, false, 2>(v, ln, lds, tr);
      const float cs = -a.scale * q2;
#pragma unroll
      for (int s = 0; s < PT; ++s) {
        const cf t = cf{fmaf(kap_x[s], v[0][s].im, kap_y * v[1][s].im),
                        -fmaf(kap_x[s], v[0][s].re, kap_y * v[1][s].re)};
        A[s] = keep_x[s] ? cs * t : cf{0.f, 0.f};
      }
    } else {
#pragma unroll
      for (int s = 0; s < PT; ++s) A[s] = cf{0.f, 0.f};
    }
  };

  if constexpr (MODE == COLA_CL) {
    cf v[1][PT];
    const cf* t = a.T + at;
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = gld<COH>(t + j + s * T);
    cfft<N, PT, false, 1>(v, ln, lds, tr);
    if (active) {
      cf* lam = aa.LAM + at;
#pragma unroll
      for (int s = 0; s < PT; ++s) lam[j + s * T] = v[0][s];
    }
    cf mu[PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) mu[s] = coef(2, s) * v[0][s];
    h_out(mu);
  } else if constexpr (MODE == COLA_CD) {
    cf A[PT], rho[PT], sg[PT];
    adj_nonlinear(A);
    const cf* lam = aa.LAM + at;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf l = lam[j + s * T];
      rho[s] = l + A[s];
      sg[s] = coef(1, s) * rho[s] - coef(2, s) * l;
    }
    if (active) {
      cf* o = aa.RHO + at;
#pragma unroll
      for (int s = 0; s < PT; ++s) o[j + s * T] = rho[s];
    }
    h_out(sg);
  } else {
    static_assert(MODE == COLA_CE, "unknown adjoint column mode");
    cf B[PT], v[1][PT];
    adj_nonlinear(B);
    const cf* rho = aa.RHO + at;
#pragma unroll
    for (int s = 0; s < PT; ++s) v[0][s] = coef(0, s) * rho[j + s * T] + B[s];
    cfft<N, PT, true, 1>(v, ln, lds, tr);
    if (active) {
      cf* t = a.T + at;
#pragma unroll
      for (int s = 0; s < PT; ++s) gst<COH>(t + j + s * T, v[0][s]);
    }
  }
}

template <int N, int MODE>
__global__ void __launch_bounds__(ColCfg<N>::WG) col2d_adj_kernel(Adj2DArgs aa) {
  __shared__ cf lds_x[ColLds<N>::X];
  col2d_adj_body<N, MODE, false>(aa, blockIdx.x, blockIdx.y, lds_x);
}

template <int N, int MODE>
inline hipError_t launch_col2d_adj(const Adj2DArgs& a, int ninst, hipStream_t st) {
  using C = ColCfg<N>;
  const int groups = (C::NY + C::CPB - 1) / C::CPB;
  hipLaunchKernelGGL((col2d_adj_kernel<N, MODE>), dim3(groups, ninst), dim3(C::WG), 0, st, a);
  return hipGetLastError();
}

}  // namespace exmi
