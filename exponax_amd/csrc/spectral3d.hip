// 3-D spectral transforms and the linear (ETDRK0) step for gfx950.
//
// rfftn over (x, y, z) of an N^3 field (_spectral.py:614-656) is a real FFT
// along z (the contiguous axis; the 1-D batched kernel of fft1d.hpp) followed
// by complex FFTs along y and then x of every line of the [x][y][kz] spectrum,
// in place; irfftn (:659-721) runs the inverse line FFTs along x and y and
// then the c2r along z. The line kernel below owns one line with T = N/PT
// lanes (lane j holds points j + s T); CPB lines share a 256-thread workgroup,
// and lines adjacent in memory sit in adjacent lane groups, so for a fixed
// slot the groups of a wavefront read neighbouring addresses.
#include "exmi_internal.hpp"
#include "fft1d.hpp"

namespace exmi {

template <int N>
struct LineCfg {
  static constexpr int PT = N <= 128 ? 4 : 8;
  static constexpr int T = N / PT;
  static constexpr int WG = 256;
  static constexpr int CPB = WG / T;  // lines per workgroup
  static_assert(T <= 64, "a line transform must fit one wavefront");
};

// Line l = (outer o, inner i) with o = l / inner, i = l % inner starts at
// o * outer_stride + i; its points are pstride apart. In place, times scale.
template <int N, bool INV>
__global__ void __launch_bounds__(256) lines_kernel(cf* data, long long nlines, long long inner, long long outer_stride,
                                                    long long pstride, float scale, const cf* __restrict__ tw) {
  using C = LineCfg<N>;
  constexpr int PT = C::PT, T = C::T;
  constexpr int LDSN = LdsSize<N>::value;
  __shared__ cf lds_x[C::CPB * LDSN];
  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const long long raw = (long long)blockIdx.x * C::CPB + c;
  const bool active = raw < nlines;
  const long long line = active ? raw : 0;  // inactive groups mirror line 0 and store nothing
  const long long base = (line / inner) * outer_stride + line % inner;
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, tw);
  cf v[1][PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) v[0][s] = data[base + (long long)(j + s * T) * pstride];
  cfft<N, PT, INV, 1>(v, ln, lds, tr);
  if (active) {
#pragma unroll
    for (int s = 0; s < PT; ++s) data[base + (long long)(j + s * T) * pstride] = scale * v[0][s];
  }
}

template <int N, bool INV>
static hipError_t launch_lines_n(cf* data, long long nlines, long long inner, long long outer_stride,
                                 long long pstride, float scale, const cf* tw, hipStream_t st) {
  using C = LineCfg<N>;
  const long long grid = (nlines + C::CPB - 1) / C::CPB;
  hipLaunchKernelGGL((lines_kernel<N, INV>), dim3((unsigned)grid), dim3(C::WG), 0, st, data, nlines, inner,
                     outer_stride, pstride, scale, tw);
  return hipGetLastError();
}

hipError_t launch_lines(int N, bool inv, cf* data, long long nlines, long long inner, long long outer_stride,
                        long long pstride, float scale, const cf* tw, hipStream_t st) {
  switch (N) {
#define EXMI_LINES(NN)                                                                                        \
  case NN:                                                                                                    \
    return inv ? launch_lines_n<NN, true>(data, nlines, inner, outer_stride, pstride, scale, tw, st)         \
               : launch_lines_n<NN, false>(data, nlines, inner, outer_stride, pstride, scale, tw, st);
    EXMI_LINES(16) EXMI_LINES(32) EXMI_LINES(64) EXMI_LINES(128) EXMI_LINES(256)
#undef EXMI_LINES
    default: return hipErrorInvalidValue;
  }
}

// ETDRK0 in spectral space (etdrk/_etdrk_0.py): X <- E X per mode, E complex
// when e_im >= 0 (its coefficient index). Instance i = channel i % C of
// trajectory i / C; tables (lin_batch, lin_channels, ncoef, nr).
__global__ void etdrk0_spec_kernel(cf* X, long long ninst, long long nr, int channels, const float* __restrict__ tables,
                                   int ncoef, int lin_batch, int lin_channels, int e_im) {
  const long long total = ninst * nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / nr, k = idx % nr;
    const long long traj = inst / channels, ch = inst % channels;
    const float* tb =
        tables + ((lin_batch > 1 ? traj : 0) * lin_channels + (lin_channels > 1 ? ch : 0)) * (long long)ncoef * nr;
    const cf x = X[idx];
    const float er = tb[k];
    if (e_im < 0) {
      X[idx] = er * x;
    } else {
      const float ei = tb[(long long)e_im * nr + k];
      X[idx] = cf{fmaf(er, x.re, -ei * x.im), fmaf(er, x.im, ei * x.re)};
    }
  }
}

hipError_t launch_etdrk0_spec(cf* X, long long ninst, long long nr, int channels, const float* tables, int ncoef,
                              int lin_batch, int lin_channels, int e_im, hipStream_t st) {
  const long long total = ninst * nr;
  long long grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(etdrk0_spec_kernel, dim3((unsigned)grid), dim3(256), 0, st, X, ninst, nr, channels, tables,
                     ncoef, lin_batch, lin_channels, e_im);
  return hipGetLastError();
}

}  // namespace exmi
