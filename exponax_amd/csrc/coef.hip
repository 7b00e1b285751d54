// ETDRK operator tables formed once per plan, in float64 on the device.
//
// For every mode: E = exp(dt L), E_half = exp(dt L / 2) and the ETDRK
// phi-coefficients as the Kassam-Trefethen contour mean over M points
// r e^{2 pi i (j - 1/2)/M} + dt L, keeping the real part of each sample and
// multiplying the mean by dt, exactly the reductions of
// etdrk/_etdrk_1.py:66-76, _etdrk_2.py:71-89, _etdrk_3.py:156-189 and
// _etdrk_4.py:170-196 (roots from etdrk/_utils.py:9-23). Results are stored
// as float32 in the layout (row, coefficient, mode). With a complex L (Li not
// null) E and E_half are complex; their imaginary parts follow the real
// coefficients (num_coefs_c), the phi-coefficients are real as above.
// With per-row factors nf (not null) every coefficient that multiplies a
// nonlinear term (all but E and E_half) is multiplied by its row's factor in
// float64 before rounding: the ETDRK stages are linear in the nonlinear
// term, so this steps N' = f N, a per-trajectory scale of the nonlinearity.
#include <hip/hip_runtime.h>

#include "exmi_internal.hpp"

namespace exmi {

struct zd {
  double re, im;
};
__device__ inline zd operator+(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline zd operator-(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline zd operator*(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline zd operator*(double s, zd a) { return {s * a.re, s * a.im}; }
__device__ inline zd operator+(double s, zd a) { return {s + a.re, a.im}; }
__device__ inline zd zdiv(zd a, zd b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
__device__ inline zd zexp(zd z) {
  double e = exp(z.re);
  double s, c;
  sincos(z.im, &s, &c);
  return {e * c, e * s};
}

__global__ void coef_kernel(const double* __restrict__ L, const double* __restrict__ Li, long long total, int nr,
                            double dt, int order, int ncirc, double radius, const double* __restrict__ nf,
                            float* __restrict__ out) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long long row = idx / nr;
  const int k = (int)(idx % nr);
  const int nco = num_coefs(order);
  float* o = out + row * (long long)num_coefs_c(order, Li != nullptr) * nr + k;
  const double l = L[idx];
  const zd ldt = {l * dt, Li ? Li[idx] * dt : 0.0};
  if (Li) {
    const zd E = zexp(ldt);
    o[0] = (float)E.re;
    o[nco * nr] = (float)E.im;
    if (order >= 3) {
      const zd Eh = zexp(0.5 * ldt);
      o[nr] = (float)Eh.re;
      o[(nco + 1) * nr] = (float)Eh.im;
    }
  } else {
    o[0] = (float)exp(l * dt);
  }
  if (order == 0) return;
  double acc[5] = {0, 0, 0, 0, 0};
  const double pi = 3.14159265358979323846;
  for (int jj = 1; jj <= ncirc; ++jj) {
    double s, c;
    sincos(2.0 * pi * (jj - 0.5) / ncirc, &s, &c);
    const zd z = zd{radius * c, radius * s} + ldt;
    const zd ez = zexp(z);
    const zd one = {1.0, 0.0};
    const zd z2 = z * z, z3 = z2 * z;
    if (order == 1 || order == 2) {
      acc[0] += zdiv(ez - one, z).re;
      if (order == 2) acc[1] += zdiv(ez - one - z, z2).re;
    } else if (order == 3) {
      const zd ezh = zexp(0.5 * z);
      acc[0] += zdiv(ezh - one, z).re;
      acc[1] += zdiv(ez - one, z).re;
      acc[2] += zdiv((-4.0 + (-1.0) * z) + ez * ((4.0 + (-3.0) * z) + z2), z3).re;
      acc[3] += zdiv(4.0 * ((2.0 + z) + ez * (-2.0 + z)), z3).re;
      acc[4] += zdiv(((-4.0 + (-3.0) * z) - z2) + ez * (4.0 + (-1.0) * z), z3).re;
    } else {
      const zd ezh = zexp(0.5 * z);
      acc[0] += zdiv(ezh - one, z).re;
      acc[1] += zdiv((-4.0 + (-1.0) * z) + ez * ((4.0 + (-3.0) * z) + z2), z3).re;
      acc[2] += zdiv((2.0 + z) + ez * (-2.0 + z), z3).re;
      acc[3] += zdiv(((-4.0 + (-3.0) * z) - z2) + ez * (4.0 + (-1.0) * z), z3).re;
    }
  }
  const double f = nf ? nf[row] : 1.0;
  if (order == 1) {
    o[nr] = (float)(f * (dt * (acc[0] / ncirc)));
  } else if (order == 2) {
    o[nr] = (float)(f * (dt * (acc[0] / ncirc)));
    o[2 * nr] = (float)(f * (dt * (acc[1] / ncirc)));
  } else if (order == 3) {
    if (!Li) o[nr] = (float)exp(0.5 * dt * l);
    for (int i = 0; i < 5; ++i) o[(2 + i) * nr] = (float)(f * (dt * (acc[i] / ncirc)));
  } else {
    if (!Li) o[nr] = (float)exp(0.5 * dt * l);
    for (int i = 0; i < 4; ++i) o[(2 + i) * nr] = (float)(f * (dt * (acc[i] / ncirc)));
  }
}

hipError_t launch_coef(const double* L, const double* Li, long long total, int nr, double dt, int order, int ncirc,
                       double radius, const double* nf, float* out, hipStream_t st) {
  const int bs = 256;
  const long long grid = (total + bs - 1) / bs;
  hipLaunchKernelGGL(coef_kernel, dim3((unsigned)grid), dim3(bs), 0, st, L, Li, total, nr, dt, order, ncirc, radius,
                     nf, out);
  return hipGetLastError();
}

}  // namespace exmi
