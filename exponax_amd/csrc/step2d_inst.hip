// Instantiates the 2-D row and column kernels (step and adjoint) for one grid size, selected at
// compile time with -DEXMI_N2=<N>. The build compiles this file once per
// supported size.
#include "adjoint2d.hpp"
#include "coop2d.hpp"
#include "roll2d.hpp"
#include "step2d.hpp"

#ifndef EXMI_N2
#error "EXMI_N2 must be defined"
#endif

namespace exmi {

using Launch2D = hipError_t (*)(const Step2DArgs&, int, hipStream_t);

#define EXMI_CAT2(a, b) a##b
#define EXMI_CAT(a, b) EXMI_CAT2(a, b)

template <int N, int KIND>
static Launch2D pick2d(int is_row, int mode) {
  if (is_row) {
    switch (mode) {
      case ROW_RA: return &launch_row2d<N, ROW_RA, KIND>;
      case ROW_RB: return &launch_row2d<N, ROW_RB, KIND>;
      case ROW_RC: return &launch_row2d<N, ROW_RC, KIND>;
      case ROW_RCB: return &launch_row2d<N, ROW_RCB, KIND>;
    }
  } else {
    switch (mode) {
      case COL_CA: return &launch_col2d<N, COL_CA, KIND>;
      case COL_CB: return &launch_col2d<N, COL_CB, KIND>;
      case COL_CC: return &launch_col2d<N, COL_CC, KIND>;
      case COL_C0: return &launch_col2d<N, COL_C0, KIND>;
      case COL_C1: return &launch_col2d<N, COL_C1, KIND>;
      case COL_C3A: return &launch_col2d<N, COL_C3A, KIND>;
      case COL_C3B: return &launch_col2d<N, COL_C3B, KIND>;
      case COL_C3C: return &launch_col2d<N, COL_C3C, KIND>;
      case COL_C4A: return &launch_col2d<N, COL_C4A, KIND>;
      case COL_C4B: return &launch_col2d<N, COL_C4B, KIND>;
      case COL_C4C: return &launch_col2d<N, COL_C4C, KIND>;
      case COL_C4D: return &launch_col2d<N, COL_C4D, KIND>;
      // the standalone transforms do not depend on the nonlinear kind
      case COL_FWD: return KIND == NL_GRADNORM ? &launch_col2d<N, COL_FWD, NL_GRADNORM> : nullptr;
      case COL_INV: return KIND == NL_GRADNORM ? &launch_col2d<N, COL_INV, NL_GRADNORM> : nullptr;
      case COL_CA_SPEC: return &launch_col2d<N, COL_CA_SPEC, KIND>;
      case COL_NOUT: return &launch_col2d<N, COL_NOUT, KIND>;
      case COL_CCA: return &launch_col2d<N, COL_CCA, KIND>;
    }
  }
  return nullptr;
}

using LaunchAdj2D = hipError_t (*)(const Adj2DArgs&, int, hipStream_t);

// adjoint modes (gradient-norm nonlinearity only)
LaunchAdj2D EXMI_CAT(get_adj2d_, EXMI_N2)(int is_row, int mode) {
  constexpr int N = EXMI_N2;
  if (is_row) {
    switch (mode) {
      case ROWA_RBG: return &launch_row2d_adj<N, ROWA_RBG>;
      case ROWA_RG: return &launch_row2d_adj<N, ROWA_RG>;
      case ROWA_RP: return &launch_row2d_adj<N, ROWA_RP>;
    }
  } else {
    switch (mode) {
      case COLA_CL: return &launch_col2d_adj<N, COLA_CL>;
      case COLA_CD: return &launch_col2d_adj<N, COLA_CD>;
      case COLA_CE: return &launch_col2d_adj<N, COLA_CE>;
    }
  }
  return nullptr;
}

Launch2D EXMI_CAT(get_2d_, EXMI_N2)(int kind, int is_row, int mode) {
  constexpr int N = EXMI_N2;
  if (kind == NL_VORT) return pick2d<N, NL_VORT>(is_row, mode);
  if (kind == NL_GRADNORM) return pick2d<N, NL_GRADNORM>(is_row, mode);
  if (kind == NL_POLY) return pick2d<N, NL_POLY>(is_row, mode);
  if (kind == NL_CONV_CONS) return pick2d<N, NL_CONV_CONS>(is_row, mode);
  if (kind == NL_CONV_NONCONS) return pick2d<N, NL_CONV_NONCONS>(is_row, mode);
  if (kind == NL_GENERAL) return pick2d<N, NL_GENERAL>(is_row, mode);
  if (kind == NL_GRAY_SCOTT) return pick2d<N, NL_GRAY_SCOTT>(is_row, mode);
  if (kind == NL_BZ) return pick2d<N, NL_BZ>(is_row, mode);
  return nullptr;
}

using LaunchRoll2D = hipError_t (*)(const Roll2DArgs&, hipStream_t);

// workgroup-resident ETDRK2 rollouts (N <= 128; gradient norm and vorticity)
LaunchRoll2D EXMI_CAT(get_roll2d_, EXMI_N2)(int kind) {
  constexpr int N = EXMI_N2;
  if constexpr (N <= 128) {
    if (kind == NL_VORT) return &launch_roll2d<N, NL_VORT>;
    if (kind == NL_GRADNORM) return &launch_roll2d<N, NL_GRADNORM>;
  }
  (void)kind;
  return nullptr;
}

using LaunchCoop2D = hipError_t (*)(const Coop2DArgs&, hipStream_t);
using Coop2DCapacity = long long (*)();

// cooperative ETDRK2 rollouts (N = 128 and 256; gradient norm and
// vorticity); *cap: instances one launch can hold
LaunchCoop2D EXMI_CAT(get_coop2d_, EXMI_N2)(int kind, Coop2DCapacity* cap) {
  constexpr int N = EXMI_N2;
  if constexpr (N == 128 || N == 256) {
    if (kind == NL_VORT) {
      *cap = &coop2d_max_instances<N, NL_VORT>;
      return &launch_coop2d<N, NL_VORT>;
    }
    if (kind == NL_GRADNORM) {
      *cap = &coop2d_max_instances<N, NL_GRADNORM>;
      return &launch_coop2d<N, NL_GRADNORM>;
    }
  }
  (void)kind;
  *cap = nullptr;
  return nullptr;
}

using LaunchCoopAdj2D = hipError_t (*)(const CoopAdj2DArgs&, hipStream_t);

// cooperative transposed segments (gradient norm; N = 128 and 256)
LaunchCoopAdj2D EXMI_CAT(get_coop_adj2d_, EXMI_N2)(Coop2DCapacity* cap) {
  constexpr int N = EXMI_N2;
  if constexpr (N == 128 || N == 256) {
    *cap = &coop_adj2d_max_instances<N>;
    return &launch_coop_adj2d<N>;
  }
  *cap = nullptr;
  return nullptr;
}

}  // namespace exmi
