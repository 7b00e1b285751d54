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
  NL_GRAY_SCOTT = 7,    // two-channel Gray-Scott reaction, channels coupled
  NL_BZ = 8,            // three-channel Belousov-Zhabotinsky reaction, channels coupled
  NL_PROJ3D = 9,        // 3-D Leray-projected rotational convection (+ Kolmogorov forcing)
  NL_NUM_KINDS = 10
};

// Channel-coupled reactions at one grid point, unit scale: Gray-Scott
// (stepper/reaction/_gray_scott.py:36-46) and Belousov-Zhabotinsky
// (_belousov_zhabotinsky.py:38-45), in the reference's operation order.
__host__ __device__ __forceinline__ void gray_scott(float u0, float u1, float f, float k, float& r0, float& r1) {
  const float uvv = u0 * u1 * u1;
  r0 = f * (1.f - u0) - uvv;
  r1 = -(f + k) * u1 + uvv;
}
__host__ __device__ __forceinline__ void belousov_zhabotinsky(float u0, float u1, float u2, float& r0, float& r1,
                                                             float& r2) {
  r0 = u0 + u1 - u0 * u1 - u0 * u0;
  r1 = u2 - u1 - u0 * u1;
  r2 = u0 - u2;
}

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
