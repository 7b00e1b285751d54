// Internal constants shared by the kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/exmi.h"

namespace exmi {

// Kernel-level nonlinear kinds (one fused evaluation routine each).
enum KernelKind {
  NL_ZERO = 0,
  NL_CONV_CONS = 1,     // conservative convection, 1 inverse + 1 forward FFT
  NL_CONV_NONCONS = 2,  // non-conservative convection, 2 inverse + 1 forward
  NL_GRADNORM = 3,      // gradient norm with zero-mode fix, 1 + 1
  NL_POLY = 4,          // polynomial reaction, 1 + 1
  NL_GENERAL = 5,       // b0 u^2 + conservative convection + gradient norm, 2 + 2
  NL_VORT = 6,          // 2-D vorticity convection (+ Kolmogorov forcing)
  NL_NUM_KINDS = 7
};

__host__ __device__ inline int num_coefs(int order) {
  switch (order) {
    case 0: return 1;
    case 1: return 2;
    case 2: return 3;
    case 3: return 7;
    case 4: return 6;
    default: return 0;
  }
}
// With a complex linear operator the imaginary parts of E (and E_half for
// orders 3 and 4) follow the real coefficients.
__host__ __device__ inline int num_coefs_c(int order, bool cplx) {
  return num_coefs(order) + (cplx ? (order >= 3 ? 2 : 1) : 0);
}

}  // namespace exmi
