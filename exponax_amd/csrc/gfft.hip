// Runtime-size mixed-radix FFTs for gfx950.
//
// The fused step kernels are compiled per transform size (powers of two and
// the two small mixed-radix sizes 48 and 50). Every other grid size the
// reference accepts (_spectral.py:614-721 takes any N: odd N, 100, 513, ...)
// runs through the kernel below: a workgroup stages `lpb` lines in LDS and
// runs the self-sorting Stockham passes of a factorisation n = R_1 ... R_m
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC 2008): in the pass of radix R after passes whose
// radices multiply to Ns, butterfly j reads x[j + s n/R] (s < R), multiplies
// input s by exp(-2 pi i s (j mod Ns) / (Ns R)), forms the R-point DFT and
// writes output q to ((j / Ns) Ns R + j mod Ns + q Ns). Each output is
// computed directly as one R-term sum whose twiddle exponent
// s ((j mod Ns) n/(Ns R) + q n/R) mod n indexes one table of the n-th roots
// of unity (formed in double precision, rounded once), in LDS when it fits.
// Radices up to 16 (and 17, 19) run as register butterflies: a thread reads
// a butterfly's R inputs once, twiddles them and forms the R-point DFT with
// the R-th roots from the same table; any other radix (a prime factor above
// the largest pass) is a direct-sum pass, so every n factors.
#include <cstdio>

#include "gfft.hpp"

namespace exmi {

constexpr int GFFT_WG = 256;

struct GfftArgs {
  GfftCfg c;
  const void* in;
  void* out;
  long long nlines, inner, outer_stride, pstride;
  float scale;
  const cf* tw;
  int tw_lds;  // 1: the twiddle table is staged in LDS (it fits beside the two line buffers)
};

// x * w, or x * conj(w) for the inverse
template <bool INV>
__device__ __forceinline__ cf twmul(cf x, cf w) {
  return INV ? cf{fmaf(x.re, w.re, x.im * w.im), fmaf(x.im, w.re, -x.re * w.im)}
             : cf{fmaf(x.re, w.re, -x.im * w.im), fmaf(x.im, w.re, x.re * w.im)};
}

// One Stockham pass of radix R (compile time) over lpb lines of n points:
// each thread owns whole butterflies, reads its R inputs once, twiddles them,
// forms the R-point DFT in registers (the R-th roots of unity are the table
// entries k n / R) and writes its R outputs.
template <int R, bool INV>
__device__ __forceinline__ void gfft_pass_r(const cf* A, cf* B, const cf* T, int n, int Ns, int lpb, int tid) {
  const int m = n / R;           // butterflies per line
  const int tws = n / (Ns * R);  // twiddle stride of this pass
  cf wr[R];
#pragma unroll
  for (int k = 0; k < R; ++k) wr[k] = T[k * m];
  for (int idx = tid; idx < lpb * m; idx += GFFT_WG) {
    const int ll = idx / m, j = idx - ll * m;
    const int jm = j % Ns;
    const cf* src = A + ll * n + j;
    cf x[R];
    x[0] = src[0];
    const int e1 = jm * tws;  // s * e1 < n for s < R
#pragma unroll
    for (int s = 1; s < R; ++s) x[s] = twmul<INV>(src[s * m], T[s * e1]);
    cf* dst = B + ll * n + (j - jm) * R + jm;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cf acc = x[0];
#pragma unroll
      for (int s = 1; s < R; ++s) {
        const cf w = wr[(q * s) % R];
        const cf t = twmul<INV>(x[s], w);
        acc.re += t.re;
        acc.im += t.im;
      }
      dst[q * Ns] = acc;
    }
  }
}

// Any radix: each output one R-term sum whose twiddle exponents
// s ((j mod Ns) n/(Ns R) + q n/R) mod n index the table.
template <bool INV>
__device__ __forceinline__ void gfft_pass_direct(const cf* A, cf* B, const cf* T, int n, int R, int Ns, int lpb,
                                                 int tid) {
  const int m = n / R;
  const int tws = n / (Ns * R);
  for (int idx = tid; idx < lpb * n; idx += GFFT_WG) {
    const int ll = idx / n, rem = idx - ll * n;
    const int q = rem / m, j = rem - q * m;
    const int jm = j % Ns;
    int es = jm * tws + q * m;
    if (es >= n) es -= n;
    const cf* src = A + ll * n + j;
    cf acc = src[0];
    int e = es;
    for (int s = 1; s < R; ++s) {
      const cf t = twmul<INV>(src[s * m], T[e]);
      acc.re += t.re;
      acc.im += t.im;
      e += es;
      if (e >= n) e -= n;
    }
    B[ll * n + (j - jm) * R + jm + q * Ns] = acc;
  }
}

// RMAX: the largest radix run as register butterflies (8 keeps the
// kernel's registers, and so its occupancy, low; 19 adds 9..17 and 19 for
// plans that have such a pass).
template <int MODE, int RMAX>
__global__ void __launch_bounds__(GFFT_WG) gfft_kernel(GfftArgs a) {
  extern __shared__ cf gfft_smem[];
  const int n = a.c.n, lpb = a.c.lpb, tot = n * lpb, nz = n / 2 + 1;
  cf* A = gfft_smem;
  cf* B = gfft_smem + tot;
  const int tid = threadIdx.x;
  const long long line0 = (long long)blockIdx.x * lpb;
  constexpr bool INV = MODE == GFFT_INV || MODE == GFFT_C2R;
  const cf* T = a.tw;
  if (a.tw_lds) {
    cf* Tl = gfft_smem + 2 * tot;
    for (int i = tid; i < n; i += GFFT_WG) Tl[i] = a.tw[i];
    T = Tl;
  }
  // load: consecutive threads read consecutive addresses (along the line for
  // contiguous lines, across neighbouring lines for a strided axis)
  const bool along = MODE != GFFT_FWD && MODE != GFFT_INV ? true : a.pstride == 1;
  for (int idx = tid; idx < tot; idx += GFFT_WG) {
    const int ll = along ? idx / n : idx % lpb;
    const int pt = along ? idx - ll * n : idx / lpb;
    const long long gl = line0 + ll;
    cf v = cf{0.f, 0.f};
    if (gl < a.nlines) {
      if constexpr (MODE == GFFT_R2C) {
        v = cf{static_cast<const float*>(a.in)[gl * n + pt], 0.f};
      } else if constexpr (MODE == GFFT_C2R) {
        const bool upper = pt > n / 2;
        const int k = upper ? n - pt : pt;
        v = static_cast<const cf*>(a.in)[gl * nz + k];
        if (k == 0 || 2 * k == n) v.im = 0.f;
        if (upper) v.im = -v.im;
      } else {
        const long long base = (gl / a.inner) * a.outer_stride + gl % a.inner;
        v = static_cast<const cf*>(a.in)[base + (long long)pt * a.pstride];
      }
    }
    A[ll * n + pt] = v;
  }
  __syncthreads();
  int Ns = 1;
  for (int p = 0; p < a.c.npass; ++p) {
    const int R = a.c.radix[p];
    switch (R) {
#define EXMI_GFFT_R(RR)                                          \
  case RR:                                                       \
    if constexpr (RR <= RMAX) {                                  \
      gfft_pass_r<RR, INV>(A, B, T, n, Ns, lpb, tid);            \
    } else {                                                     \
      gfft_pass_direct<INV>(A, B, T, n, R, Ns, lpb, tid);        \
    }                                                            \
    break;
      EXMI_GFFT_R(2) EXMI_GFFT_R(3) EXMI_GFFT_R(4) EXMI_GFFT_R(5) EXMI_GFFT_R(6) EXMI_GFFT_R(7)
      EXMI_GFFT_R(8) EXMI_GFFT_R(9) EXMI_GFFT_R(10) EXMI_GFFT_R(11) EXMI_GFFT_R(12) EXMI_GFFT_R(13)
      EXMI_GFFT_R(14) EXMI_GFFT_R(15) EXMI_GFFT_R(16) EXMI_GFFT_R(17) EXMI_GFFT_R(19)
#undef EXMI_GFFT_R
      default: gfft_pass_direct<INV>(A, B, T, n, R, Ns, lpb, tid); break;
    }
    __syncthreads();
    cf* t = A;
    A = B;
    B = t;
    Ns *= R;
  }
  // store
  const int nout = MODE == GFFT_R2C ? nz : n;
  for (int idx = tid; idx < lpb * nout; idx += GFFT_WG) {
    const int ll = along ? idx / nout : idx % lpb;
    const int pt = along ? idx - ll * nout : idx / lpb;
    const long long gl = line0 + ll;
    if (gl >= a.nlines) continue;
    const cf v = A[ll * n + pt];
    if constexpr (MODE == GFFT_R2C) {
      static_cast<cf*>(a.out)[gl * nz + pt] = v;
    } else if constexpr (MODE == GFFT_C2R) {
      static_cast<float*>(a.out)[gl * n + pt] = a.scale * v.re;
    } else {
      const long long base = (gl / a.inner) * a.outer_stride + gl % a.inner;
      static_cast<cf*>(a.out)[base + (long long)pt * a.pstride] = a.scale * v;
    }
  }
}

static int smallest_factor(int n) {
  for (int f = 2; f * f <= n; ++f)
    if (n % f == 0) return f;
  return n;
}

bool gfft_make_cfg(int n, int max_radix, int points_per_wg, GfftCfg* cfg) {
  if (n < 2 || n > GFFT_MAX_N || max_radix < 2) return false;
  GfftCfg c;
  c.n = n;
  // prime factors in ascending order, merged greedily into passes of at most
  // max_radix points (a larger prime is a pass of its own)
  int rest = n, acc = 1;
  while (rest > 1) {
    const int f = smallest_factor(rest);
    rest /= f;
    if (acc > 1 && acc * f > max_radix) {
      if (c.npass >= GFFT_MAX_PASS) return false;
      c.radix[c.npass++] = acc;
      acc = 1;
    }
    acc *= f;
  }
  if (acc > 1) {
    if (c.npass >= GFFT_MAX_PASS) return false;
    c.radix[c.npass++] = acc;
  }
  const int lpb = points_per_wg / n;
  c.lpb = lpb < 1 ? 1 : lpb;
  if (c.lpb * n > GFFT_MAX_N) c.lpb = 1;  // two line buffers stay within 64 KiB of LDS
  *cfg = c;
  return true;
}

static const int k_gfft_cands[] = {4 * 100000 + 1024,  8 * 100000 + 1024,  16 * 100000 + 1024,
                                   4 * 100000 + 2048,  8 * 100000 + 2048,  16 * 100000 + 2048,
                                   8 * 100000 + 512,   8 * 100000 + 4096,  100000000 + 16 * 100000 + 1024,
                                   100000000 + 16 * 100000 + 2048, 100000000 + 8 * 100000 + 2048};
int gfft_num_candidates() { return (int)(sizeof k_gfft_cands / sizeof k_gfft_cands[0]); }
int gfft_candidate(int i) { return k_gfft_cands[i]; }
int gfft_default_code() { return 8 * 100000 + 2048; }
bool gfft_cfg_from_code(int n, int code, GfftCfg* cfg) {
  const int flag = code / 100000000, max_radix = (code / 100000) % 1000;
  if (code <= 100000 || flag > 1 || !gfft_make_cfg(n, max_radix, code % 100000, cfg)) return false;
  cfg->reg19 = flag;
  return true;
}

hipError_t launch_gfft(const GfftCfg& c, int mode, const void* in, void* out, long long nlines, long long inner,
                       long long outer_stride, long long pstride, float scale, const cf* tw, hipStream_t st) {
  if (nlines <= 0) return hipSuccess;
  if (c.n < 2 || c.n > GFFT_MAX_N || c.lpb < 1 || inner < 1) return hipErrorInvalidValue;
  GfftArgs a;
  a.c = c;
  a.in = in;
  a.out = out;
  a.nlines = nlines;
  a.inner = inner;
  a.outer_stride = outer_stride;
  a.pstride = pstride;
  a.scale = scale;
  a.tw = tw;
  const long long grid = (nlines + c.lpb - 1) / c.lpb;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  if ((long long)c.lpb * c.n > GFFT_MAX_N && c.lpb > 1) return hipErrorInvalidValue;
  size_t lds = sizeof(cf) * (size_t)(2 * c.lpb * c.n + c.n);
  a.tw_lds = lds <= 65536;
  if (!a.tw_lds) lds = sizeof(cf) * (size_t)(2 * c.lpb * c.n);
  bool big = false;  // a pass of radix 9..17 or 19 to run as register butterflies
  for (int i = 0; i < c.npass; ++i) big = big || (c.radix[i] > 8 && (c.radix[i] <= 17 || c.radix[i] == 19));
  big = big && c.reg19;
#define EXMI_GFFT_LAUNCH(MODE_)                                                                                 \
  if (big)                                                                                                      \
    hipLaunchKernelGGL((gfft_kernel<MODE_, 19>), dim3((unsigned)grid), dim3(GFFT_WG), lds, st, a);             \
  else                                                                                                          \
    hipLaunchKernelGGL((gfft_kernel<MODE_, 8>), dim3((unsigned)grid), dim3(GFFT_WG), lds, st, a);
  switch (mode) {
    case GFFT_FWD: EXMI_GFFT_LAUNCH(GFFT_FWD) break;
    case GFFT_INV: EXMI_GFFT_LAUNCH(GFFT_INV) break;
    case GFFT_R2C: EXMI_GFFT_LAUNCH(GFFT_R2C) break;
    case GFFT_C2R: EXMI_GFFT_LAUNCH(GFFT_C2R) break;
    default: return hipErrorInvalidValue;
  }
#undef EXMI_GFFT_LAUNCH
  return hipGetLastError();
}

}  // namespace exmi
