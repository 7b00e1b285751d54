// Launch arguments of the 3-D nonlinear and ETDRK stage kernels
// (spectral3d.hip), shared with the C-ABI layer.
#pragma once
#include "exmi_internal.hpp"

namespace exmi {

// A grid of any dimension D <= 3 is held as three axes (nx, ny, N) with the
// real transform along the last one: D = 3 is (N, N, N), D = 2 (1, N, N),
// D = 1 (1, 1, N). Spatial derivative d < D acts along axis 3 - D + d.
struct NL3DArgs {
  int N, NZ, channels, kind, kmax, nf, np;
  int dims, nx, ny;
  int nh, nq;       // adjoint: cotangent fields and adjoint products per instance
  long long Nr, N3, ninst;
  float kscale, scale, g0, g1, g2, feed, kill;
  int zero_mode_fix, npoly, poly_lap;
  int prod_nc;      // general kind: first product u sum_d d_d u instead of u^2
  float poly[EXMI_MAX_POLY];
  int inj_mode;     // Kolmogorov forcing of channel 0 at mode inj_mode of axis inj_axis, 0 on the others; -1: none
  int inj_axis;     // 1 (3-D projected convection, the y axis) or 2 (2-D vorticity, the last axis)
  float inj_value;  // injection_scale * coef_extraction scaling
};

// Which nonlinear kinds the multi-pass adjoint covers (every order): the
// gradient norm, both multi-channel convections, the polynomial (not
// Cahn-Hilliard), the general form (single-channel convections included)
// and the 2-D vorticity convection; 0 cotangent fields for others.
int adj3d_fields(int kind, int dims, int poly_lap, int* nh, int* nq);

// out = sum_t mult_t * op_t(x_t) on whole spectra, op: -1 identity,
// 100 conj(E), 101 conj(E_half), 0 <= i < ncoef the real coefficient i of
// the tables (the host checks the range).
struct AdjLinArgs {
  long long ninst, nr;
  int channels, ncoef, lin_batch, lin_channels, e_im, nt;
  const float* tables;
  const cf* x[4];
  int op[4];
  float mult[4];
  cf* out;
};

// The stage combinations of etdrk/_etdrk_{1,2,3,4}.py on whole spectra.
enum Stage3D { ST1 = 0, ST2A, ST2B, ST3A, ST3B, ST3C, ST4A, ST4B, ST4C, ST4D };

struct Stage3DArgs {
  long long ninst, nr;
  int channels, ncoef, lin_batch, lin_channels, e_im, mode;
  const float* tables;
  const cf* U0;
  const cf* N0;
  const cf* NC;
  cf* S;
  cf* S1;
  cf* A;
};

}  // namespace exmi
