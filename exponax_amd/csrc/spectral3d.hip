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
#include <type_traits>

#include "exmi_internal.hpp"
#include "fft1d.hpp"
#include "spectral3d.hpp"

namespace exmi {

// 16 points per lane up to N = 128 (T = N/16 lanes per line): a wavefront
// then holds 64/T neighbouring lines, and for a fixed slot it reads 64/T
// adjacent entries (128 bytes at N = 64) instead of four; 10% faster 3-D
// steps than 4 points per lane (Navier-Stokes velocity 64^3).
template <int N>
struct LineCfg {
  static constexpr int PT = N <= 128 ? 16 : 8;
  static constexpr int T = N / PT;
  static constexpr int WG = 256;
  static constexpr int CPB = WG / T;  // lines per workgroup
  static_assert(T <= 64, "a line transform must fit one wavefront");
};

// The fused x-line kernels that carry several lines per lane group or a
// stage (xstage*, fields_unit_lines): 8 points per lane from N = 64 on,
// twice the wavefronts of the line kernels' 16.
template <int N>
struct XStageCfg {
  static constexpr int PT = N >= 64 ? 8 : LineCfg<N>::PT;
  static constexpr int T = N / PT;
  static constexpr int CPB = 256 / T;
};

// Line l = (outer o, inner i) with o = l / inner, i = l % inner starts at
// o * outer_stride + i; its points are pstride apart. Read from src, written
// to data at the same offsets (in place when src == data), times scale.
// With kkeep < nz (inner a multiple of nz, a line's last-axis index l % nz),
// only the lines whose last-axis index is below kkeep run (nlines counts
// all lines): the kz planes above a dealias cutoff, zero on the way in and
// discarded on the way out.
template <int N, bool INV>
__global__ void __launch_bounds__(256) lines_kernel(const cf* src, cf* data, long long nlines, long long inner,
                                                    long long outer_stride, long long pstride, float scale,
                                                    const cf* __restrict__ tw, int nz, int kkeep) {
  using C = LineCfg<N>;
  constexpr int PT = C::PT, T = C::T;
  constexpr int LDSN = LdsSize<N>::value;
  __shared__ cf lds_x[C::CPB * LDSN];
  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const long long raw = (long long)blockIdx.x * C::CPB + c;
  const long long nrun = nlines / nz * kkeep;
  const bool active = raw < nrun;
  const long long run = active ? raw : 0;  // inactive groups mirror line 0 and store nothing
  const long long line = kkeep == nz ? run : (run / kkeep) * nz + run % kkeep;
  const long long base = (line / inner) * outer_stride + line % inner;
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, tw);
  cf v[1][PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) v[0][s] = src[base + (long long)(j + s * T) * pstride];
  cfft<N, PT, INV, 1>(v, ln, lds, tr);
  if (active) {
#pragma unroll
    for (int s = 0; s < PT; ++s) data[base + (long long)(j + s * T) * pstride] = scale * v[0][s];
  }
}

template <int N, bool INV>
static hipError_t launch_lines_n(const cf* src, cf* data, long long nlines, long long inner, long long outer_stride,
                                 long long pstride, float scale, const cf* tw, int nz, int kkeep, hipStream_t st) {
  using C = LineCfg<N>;
  const long long grid = (nlines / nz * kkeep + C::CPB - 1) / C::CPB;
  hipLaunchKernelGGL((lines_kernel<N, INV>), dim3((unsigned)grid), dim3(C::WG), 0, st, src, data, nlines, inner,
                     outer_stride, pstride, scale, tw, nz, kkeep);
  return hipGetLastError();
}

hipError_t launch_lines(int N, bool inv, cf* data, long long nlines, long long inner, long long outer_stride,
                        long long pstride, float scale, const cf* tw, hipStream_t st, const cf* src, int kkeep) {
  if (!src) src = data;
  const int nz = N / 2 + 1;
  if (kkeep <= 0 || kkeep > nz) kkeep = nz;
  if (inner % nz != 0 || nlines % nz != 0) kkeep = nz;
  switch (N) {
#define EXMI_LINES(NN)                                                                                        \
  case NN:                                                                                                    \
    return inv ? launch_lines_n<NN, true>(src, data, nlines, inner, outer_stride, pstride, scale, tw, nz,    \
                                          kkeep, st)                                                          \
               : launch_lines_n<NN, false>(src, data, nlines, inner, outer_stride, pstride, scale, tw, nz,   \
                                           kkeep, st);
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


// ------------------------------------------------------ 3-D nonlinear terms
// One evaluation N(S) of a batch of 3-D spectra (instance i = channel i % C
// of trajectory i / C) in four passes around the transforms:
//   fields   : the dealiased spectral fields of each instance (the state and
//              its derivatives i kappa_d S)        -> F (NF per instance)
//   (irfftn of every field)                        -> R
//   products : the pointwise products, reading the fields of the
//              trajectory's other channels where the kind couples them
//                                                  -> PR (NP per instance)
//   (rfftn of every product)                       -> F
//   combine  : derivative multiplies, scales, the post-dealias -> N
// The kinds and their sign conventions follow the 1-D and 2-D kernels:
// polynomial (nonlin_fun/_polynomial.py; with the Laplacian for
// Cahn-Hilliard), gradient norm (_gradient_norm.py), the three-channel
// convections (_convection.py:104-171), the general form
// (_general_nonlinear.py) and the coupled reactions.
struct Mode3D {
  float kx, ky, kz;  // kappa = (2 pi / L) k along the three axes (0 on absent ones)
  bool keep;         // |k| <= kmax on every axis
  bool zero;         // the mean mode
  bool inj;          // the Kolmogorov forcing mode
};

// Signed wavenumber of index i on a full (fftfreq) axis of n points: i for
// i <= (n - 1) / 2, else i - n (even n: the Nyquist index n/2 is -n/2).
__device__ __forceinline__ int signed_k(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

__device__ __forceinline__ Mode3D mode3d_at(const NL3DArgs& a, int ix, int iy, int iz) {
  const int kx = signed_k(ix, a.nx), ky = signed_k(iy, a.ny);
  Mode3D m;
  m.kx = a.kscale * (float)kx;
  m.ky = a.kscale * (float)ky;
  m.kz = a.kscale * (float)iz;
  m.keep = (kx < 0 ? -kx : kx) <= a.kmax && (ky < 0 ? -ky : ky) <= a.kmax && iz <= a.kmax;
  m.zero = kx == 0 && ky == 0 && iz == 0;
  m.inj = a.inj_axis == 2 ? (kx == 0 && ky == 0 && iz == a.inj_mode) : (kx == 0 && ky == a.inj_mode && iz == 0);
  return m;
}

__device__ __forceinline__ Mode3D mode3d(const NL3DArgs& a, long long k) {
  const int iz = (int)(k % a.NZ);
  const long long r = k / a.NZ;
  return mode3d_at(a, (int)(r / a.ny), (int)(r % a.ny), iz);
}

// kappa of spatial derivative d (axis 3 - D + d)
__device__ __forceinline__ float kap_d(const NL3DArgs& a, const Mode3D& m, int d) {
  const int axis = 3 - a.dims + d;
  return axis == 0 ? m.kx : axis == 1 ? m.ky : m.kz;
}

__device__ __forceinline__ cf imul(float kap, cf w) { return cf{-kap * w.im, kap * w.re}; }

// Field f of instance inst at entry k (mode m) of the spectra S.
__device__ __forceinline__ cf field_at(const NL3DArgs& a, const cf* __restrict__ S, long long inst, int f,
                                       long long k, const Mode3D& m) {
  if (!m.keep) return cf{0.f, 0.f};
  if (a.kind == NL_PROJ3D && f == 1) {
    // the curl component omega_c = (i kappa x u_hat)_c of the dealiased
    // spectra (_projected_convection.py:8-16, 120-123); field 0 is u_c
    const long long b = (inst - inst % a.channels) * a.Nr + k;
    const int c = (int)(inst % a.channels);
    const cf w0 = S[b], w1 = S[b + a.Nr], w2 = S[b + 2 * a.Nr];
    return c == 0 ? imul(m.ky, w2) - imul(m.kz, w1) : c == 1 ? imul(m.kz, w0) - imul(m.kx, w2)
                                                           : imul(m.kx, w1) - imul(m.ky, w0);
  }
  const cf w = S[inst * a.Nr + k];
  if (a.kind == NL_GRADNORM) return imul(kap_d(a, m, f), w);
  if (a.kind == NL_CONV_NONCONS || a.kind == NL_GENERAL) return f == 0 ? w : imul(kap_d(a, m, f - 1), w);
  if (a.kind == NL_VORT) {
    // 2-D (_vorticity_convection.py:78-99): psi = Laplace^-1 omega (1 at
    // the mean mode), u = d_y psi, v = -d_x psi, omega_x, omega_y; x is
    // axis 1, y the last axis
    const float kx = m.ky, ky = m.kz;
    const float lap = -(kx * kx + ky * ky);
    const cf psi = (lap == 0.f ? 1.f : 1.f / lap) * w;
    return f == 0 ? imul(ky, psi) : f == 1 ? imul(kx, w) : f == 2 ? imul(-kx, psi) : imul(ky, w);
  }
  return w;
}

__global__ void fields3d_kernel(NL3DArgs a, const cf* __restrict__ S, cf* __restrict__ F) {
  const long long total = a.ninst * a.Nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / a.Nr, k = idx % a.Nr;
    const Mode3D m = mode3d(a, k);
    cf* f = F + inst * a.nf * a.Nr + k;
    for (int q = 0; q < a.nf; ++q) f[q * a.Nr] = field_at(a, S, inst, q, k, m);
  }
}

// The fields pass fused with the first inverse transform pass of a 3-D grid
// (the x lines): line l = (field g = inst * nf + f, iy, iz) reads the state
// spectra S along x, forms field f of instance inst at every point, and
// writes its inverse FFT (times scale) to F. Lines adjacent in memory sit in
// adjacent lane groups, as in lines_kernel. Only the lines with iz < kkeep
// run (the fields are zero above the dealias cutoff; the later passes read
// those planes as zero or skip them).
template <int N>
__global__ void __launch_bounds__(256) fields_lines_kernel(NL3DArgs a, const cf* __restrict__ S, cf* __restrict__ F,
                                                           float scale, const cf* __restrict__ tw, int kkeep) {
  using C = LineCfg<N>;
  constexpr int PT = C::PT, T = C::T;
  constexpr int LDSN = LdsSize<N>::value;
  __shared__ cf lds_x[C::CPB * LDSN];
  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const long long inner = (long long)N * a.NZ;  // lines per field
  const long long nrun = a.ninst * a.nf * N * kkeep;
  const long long raw = (long long)blockIdx.x * C::CPB + c;
  const bool active = raw < nrun;
  const long long run = active ? raw : 0;  // inactive groups mirror line 0 and store nothing
  const long long line = (run / kkeep) * a.NZ + run % kkeep;
  const long long g = line / inner, li = line % inner;
  const long long inst = g / a.nf;
  const int f = (int)(g % a.nf);
  const int iy = (int)(li / a.NZ), iz = (int)(li % a.NZ);
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, tw);
  cf v[1][PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int ix = j + s * T;
    v[0][s] = field_at(a, S, inst, f, (long long)ix * inner + li, mode3d_at(a, ix, iy, iz));
  }
  cfft<N, PT, true, 1>(v, ln, lds, tr);
  if (active) {
    cf* o = F + g * a.Nr + li;
#pragma unroll
    for (int s = 0; s < PT; ++s) o[(long long)(j + s * T) * inner] = scale * v[0][s];
  }
}

__global__ void products3d_kernel(NL3DArgs a, const float* __restrict__ R, float* __restrict__ PR) {
  const long long total = a.ninst * a.N3;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / a.N3, p = idx % a.N3;
    const long long base = inst - inst % a.channels;
    const int ch = (int)(inst % a.channels);
    auto fld = [&](long long i, int f) { return R[(i * a.nf + f) * a.N3 + p]; };
    float* out = PR + inst * a.np * a.N3 + p;
    switch (a.kind) {
      case NL_POLY: {
        const float u = fld(inst, 0);
        float pw = 1.f, acc = 0.f;
        for (int i = 0; i < a.npoly; ++i) {
          acc += a.poly[i] * pw;
          pw *= u;
        }
        out[0] = acc;
        break;
      }
      case NL_GRADNORM: {
        float acc = 0.f;
        for (int d = 0; d < a.dims; ++d) {
          const float g = fld(inst, d);
          acc = fmaf(g, g, acc);
        }
        out[0] = acc;
        break;
      }
      case NL_CONV_CONS: {
        const float uc = fld(inst, 0);
        for (int j = 0; j < a.dims; ++j) out[(long long)j * a.N3] = uc * fld(base + j, 0);
        break;
      }
      case NL_CONV_NONCONS: {
        float acc = 0.f;
        for (int j = 0; j < a.dims; ++j) acc += fld(base + j, 0) * fld(inst, 1 + j);
        out[0] = acc;
        break;
      }
      case NL_GENERAL: {
        const float u = fld(inst, 0);
        float gs = 0.f, g2 = 0.f;
        for (int d = 0; d < a.dims; ++d) {
          const float g = fld(inst, 1 + d);
          gs += g;
          g2 = fmaf(g, g, g2);
        }
        // single-channel non-conservative convection (nonlin_fun/_convection.py:204-231)
        out[0] = a.prod_nc ? u * gs : u * u;
        out[a.N3] = g2;
        break;
      }
      case NL_VORT: out[0] = fld(inst, 0) * fld(inst, 1) + fld(inst, 2) * fld(inst, 3); break;
      case NL_GRAY_SCOTT: {
        float r0, r1;
        gray_scott(fld(base, 0), fld(base + 1, 0), a.feed, a.kill, r0, r1);
        out[0] = ch == 0 ? r0 : r1;
        break;
      }
      case NL_BZ: {
        float r0, r1, r2;
        belousov_zhabotinsky(fld(base, 0), fld(base + 1, 0), fld(base + 2, 0), r0, r1, r2);
        out[0] = ch == 0 ? r0 : ch == 1 ? r1 : r2;
        break;
      }
      case NL_PROJ3D: {
        // (u x omega)_c
        const int c1 = (ch + 1) % 3, c2 = (ch + 2) % 3;
        out[0] = fld(base + c1, 0) * fld(base + c2, 1) - fld(base + c2, 0) * fld(base + c1, 1);
        break;
      }
      default: out[0] = 0.f;
    }
  }
}

// N_hat at entry idx (instance idx / Nr) from the product spectra PH.
__device__ __forceinline__ cf combine_at(const NL3DArgs& a, const cf* __restrict__ PH, long long idx) {
  {
    const long long inst = idx / a.Nr, k = idx % a.Nr;
    const Mode3D m = mode3d(a, k);
    const cf* ph = PH + inst * a.np * a.Nr + k;
    cf r = cf{0.f, 0.f};
    switch (a.kind) {
      case NL_POLY:
        r = a.poly_lap ? (a.scale * -(m.kx * m.kx + m.ky * m.ky + m.kz * m.kz)) * ph[0] : ph[0];
        break;
      case NL_GRADNORM: r = (a.zero_mode_fix && m.zero) ? cf{0.f, 0.f} : (-0.5f * a.scale) * ph[0]; break;
      case NL_CONV_CONS: {
        cf acc = cf{0.f, 0.f};
        for (int j = 0; j < a.dims; ++j) acc = acc + imul(kap_d(a, m, j), ph[j * a.Nr]);
        r = (-0.5f * a.scale) * acc;
        break;
      }
      case NL_CONV_NONCONS:
      case NL_VORT: r = -a.scale * ph[0]; break;
      case NL_GENERAL: {
        const float cg = (a.zero_mode_fix && m.zero) ? 0.f : 0.5f * a.g2;
        r = a.g0 * ph[0] + imul(0.5f * a.g1 * (m.kx + m.ky + m.kz), ph[0]) + cg * ph[a.Nr];
        break;
      }
      case NL_PROJ3D: {
        // Leray projection (_leray.py:86-103): c + i kappa_c p with
        // p = -div / Laplace = (i kappa . c) / |kappa|^2 (p = 0 at the mean mode)
        const long long b = (inst - inst % a.channels) * a.Nr + k;  // np = 1
        const cf p0 = PH[b], p1 = PH[b + a.Nr], p2 = PH[b + 2 * a.Nr];
        const cf div = imul(m.kx, p0) + imul(m.ky, p1) + imul(m.kz, p2);
        const float k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
        const cf pr = k2 != 0.f ? (1.f / k2) * div : cf{0.f, 0.f};
        const int c = (int)(inst % a.channels);
        r = ph[0] + imul(c == 0 ? m.kx : c == 1 ? m.ky : m.kz, pr);
        break;
      }
      default: r = ph[0];
    }
    cf out = m.keep ? r : cf{0.f, 0.f};
    if ((a.kind == NL_PROJ3D || a.kind == NL_VORT) && m.inj && inst % a.channels == 0)
      out = out + cf{a.inj_value, 0.f};
    return out;
  }
}

__global__ void combine3d_kernel(NL3DArgs a, const cf* __restrict__ PH, cf* __restrict__ Nh) {
  const long long total = a.ninst * a.Nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x)
    Nh[idx] = combine_at(a, PH, idx);
}

// The z rows of the projected convection (Navier-Stokes velocity, three
// channels), fused: one group of T lanes owns the row (trajectory, ix, iy)
// of the six fields (u_c, omega_c) after their inverse x and y lines, runs
// their six c2r transforms along z (the planes above kcut read as zero),
// forms the cross products (u x omega)_c = u_c1 omega_c2 - u_c2 omega_c1
// (c1 = c + 1, c2 = c + 2 mod 3; _projected_convection.py:120-123) at every
// point in registers, and stores their r2c transforms along z for kz <= kcut
// to P (instance traj * 3 + c, the layout of the spectra). Replaces the z
// c2r pass, the products pass and the z r2c pass, and the physical fields'
// round trip through memory.
template <int N, int PT>
__global__ void __launch_bounds__(256) zcross3d_kernel(const cf* __restrict__ F, cf* __restrict__ P, long long nrows,
                                                       long long Nr, int kcut, RealTables rt) {
  constexpr int M = N / 2, T = M / PT, G = 256 / T, NZ = M + 1;
  constexpr int LDSN = LdsSize<M>::value;
  __shared__ cf lds_all[G * LDSN];
  const int g = threadIdx.x / T, j = threadIdx.x % T;
  const long long raw = (long long)blockIdx.x * G + g;
  const bool active = raw < nrows;
  const long long row = active ? raw : 0;  // inactive groups mirror row 0 and store nothing
  const long long traj = row / ((long long)N * N), rr = row % ((long long)N * N);
  cf* lds = lds_all + g * LDSN;
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> tr;
  tr.load(j, rt);
  cf v[6][PT];
  cf nyq[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const cf* x = F + (traj * 6 + q) * Nr + rr * NZ;  // field q = 2 c + f of instance traj * 3 + c
#pragma unroll
    for (int s = 0; s < PT; ++s) v[q][s] = j + s * T <= kcut ? x[j + s * T] : cf{0.f, 0.f};
    nyq[q] = M <= kcut ? x[M] : cf{0.f, 0.f};
  }
  irfft<M, PT, 6>(v, nyq, ln, lds, tr);  // N times the physical pairs
  constexpr float rn2 = 1.f / ((float)N * (float)N);
  cf pr[3][PT];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf u1 = v[2 * c1][s], w2 = v[2 * c2 + 1][s], u2 = v[2 * c2][s], w1 = v[2 * c1 + 1][s];
      pr[c][s] = cf{(u1.re * w2.re - u2.re * w1.re) * rn2, (u1.im * w2.im - u2.im * w1.im) * rn2};
    }
  }
  cf pn[3];
  rfft<M, PT, 3>(pr, pn, ln, lds, tr, kcut);  // 2 X
  if (active) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      cf* y = P + (traj * 3 + c) * Nr + rr * NZ;
#pragma unroll
      for (int s = 0; s < PT; ++s)
        if (j + s * T <= kcut) y[j + s * T] = 0.5f * pr[c][s];
      if (j == 0 && M <= kcut) y[M] = 0.5f * pn[c];
    }
  }
}

// The z rows of the other nonlinear kinds, fused the same way (the c2r of
// the trajectory's fields, the products of products3d_kernel at every point
// in registers, the r2c of the products): C instances per trajectory, NF
// fields and NP products per instance. Gradient norm (3 fields, 1 product),
// the general form (u and its 3 derivatives; u^2 or u sum_d d_d u, and
// |grad u|^2), the polynomial and the coupled reactions (one field and one
// product per channel), the three-channel convections (conservative: u_c,
// and u_c u_j for every j; non-conservative: u_c and its 3 derivatives, and
// sum_j u_j d_j u_c).
template <int KIND>
struct ZProd {
  static constexpr bool CONV = KIND == NL_CONV_CONS || KIND == NL_CONV_NONCONS;
  static constexpr int C = KIND == NL_GRAY_SCOTT ? 2 : (KIND == NL_BZ || CONV || KIND == NL_PROJ3D) ? 3 : 1;
  static constexpr int NF = KIND == NL_GRADNORM ? 3
                            : (KIND == NL_GENERAL || KIND == NL_CONV_NONCONS) ? 4
                            : KIND == NL_PROJ3D ? 2
                                                : 1;
  static constexpr int NP = KIND == NL_GENERAL ? 2 : KIND == NL_CONV_CONS ? 3 : 1;
};

// The products at one physical point from the trajectory's fields f[q]
// (q = c NF + field of instance c) into o[q] (q = c NP + product); the
// formulas of products3d_kernel.
template <int KIND>
__device__ __forceinline__ void zpoint(const NL3DArgs& a, const float* f, float* o) {
  if constexpr (KIND == NL_GRADNORM) {
    o[0] = fmaf(f[2], f[2], fmaf(f[1], f[1], f[0] * f[0]));
  } else if constexpr (KIND == NL_GENERAL) {
    const float gs = f[1] + f[2] + f[3];
    o[0] = a.prod_nc ? f[0] * gs : f[0] * f[0];
    o[1] = fmaf(f[3], f[3], fmaf(f[2], f[2], f[1] * f[1]));
  } else if constexpr (KIND == NL_POLY) {
    float pw = 1.f, acc = 0.f;
    for (int i = 0; i < a.npoly; ++i) {
      acc += a.poly[i] * pw;
      pw *= f[0];
    }
    o[0] = acc;
  } else if constexpr (KIND == NL_GRAY_SCOTT) {
    gray_scott(f[0], f[1], a.feed, a.kill, o[0], o[1]);
  } else if constexpr (KIND == NL_CONV_CONS) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) o[c * 3 + jj] = f[c] * f[jj];
  } else if constexpr (KIND == NL_CONV_NONCONS) {
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = fmaf(f[8], f[c * 4 + 3], fmaf(f[4], f[c * 4 + 2], f[0] * f[c * 4 + 1]));
  } else if constexpr (KIND == NL_PROJ3D) {
    // (u x omega)_c = u_c1 omega_c2 - u_c2 omega_c1 (_projected_convection.py:120-123)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      o[c] = f[2 * c1] * f[2 * c2 + 1] - f[2 * c2] * f[2 * c1 + 1];
    }
  } else {
    belousov_zhabotinsky(f[0], f[1], f[2], o[0], o[1], o[2]);
  }
}

template <int N, int PT, int KIND>
__global__ void __launch_bounds__(256) zprod3d_kernel(NL3DArgs a, const cf* __restrict__ F, cf* __restrict__ P,
                                                      long long nrows, int kcut, RealTables rt) {
  using Z = ZProd<KIND>;
  constexpr int C = Z::C, NF = Z::NF, NP = Z::NP, NIN = C * NF, NOUT = C * NP;
  constexpr int M = N / 2, T = M / PT, G = 256 / T, NZ = M + 1;
  constexpr int LDSN = LdsSize<M>::value;
  __shared__ cf lds_all[G * LDSN];
  const int g = threadIdx.x / T, j = threadIdx.x % T;
  const long long raw = (long long)blockIdx.x * G + g;
  const bool active = raw < nrows;
  const long long row = active ? raw : 0;  // inactive groups mirror row 0 and store nothing
  const long long traj = row / ((long long)N * N), rr = row % ((long long)N * N);
  cf* lds = lds_all + g * LDSN;
  const Lane<M, PT> ln(j);
  TwRegs<M, PT> tr;
  tr.load(j, rt);
  cf v[NIN][PT];
  cf nyq[NIN];
#pragma unroll
  for (int q = 0; q < NIN; ++q) {
    const cf* x = F + (traj * NIN + q) * a.Nr + rr * NZ;  // field q % NF of instance traj * C + q / NF
#pragma unroll
    for (int s = 0; s < PT; ++s) v[q][s] = j + s * T <= kcut ? x[j + s * T] : cf{0.f, 0.f};
    nyq[q] = M <= kcut ? x[M] : cf{0.f, 0.f};
  }
  irfft<M, PT, NIN>(v, nyq, ln, lds, tr);  // N times the physical pairs
  constexpr float rn = 1.f / (float)N;
  cf pr[NOUT][PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // the two real points of the slot
      float f[NIN], o[NOUT];
#pragma unroll
      for (int q = 0; q < NIN; ++q) f[q] = (h ? v[q][s].im : v[q][s].re) * rn;
      zpoint<KIND>(a, f, o);
#pragma unroll
      for (int q = 0; q < NOUT; ++q) {
        if (h) pr[q][s].im = o[q];
        else pr[q][s].re = o[q];
      }
    }
  }
  cf pn[NOUT];
  rfft<M, PT, NOUT>(pr, pn, ln, lds, tr, kcut);  // 2 X
  if (active) {
#pragma unroll
    for (int q = 0; q < NOUT; ++q) {
      cf* y = P + (traj * NOUT + q) * a.Nr + rr * NZ;  // product q % NP of instance traj * C + q / NP
#pragma unroll
      for (int s = 0; s < PT; ++s)
        if (j + s * T <= kcut) y[j + s * T] = 0.5f * pr[q][s];
      if (j == 0 && M <= kcut) y[M] = 0.5f * pn[q];
    }
  }
}

// Fused z rows of a 3-D kind other than the projected convection; returns
// hipErrorInvalidValue for the kinds it does not cover (the caller then runs
// the separate passes). P may equal F when every instance has one field and
// one product (each row group reads its rows before it writes them).
hipError_t launch_zprod3d(const NL3DArgs& a, const cf* F, cf* P, int kcut, const RealTables& rt, hipStream_t st) {
  const int N = a.N;
  auto go = [&](auto kern, int G) {
    const int C = a.kind == NL_GRAY_SCOTT ? 2 : (a.kind == NL_BZ || a.kind == NL_CONV_CONS || a.kind == NL_CONV_NONCONS) ? 3 : 1;
    const long long nrows = a.ninst / C * N * N;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nrows + G - 1) / G)), dim3(256), 0, st, a, F, P, nrows, kcut, rt);
    return hipGetLastError();
  };
  // 8 points per lane (TT lanes per row); the convections, with 12 values
  // per point in flight, 4 (2 TT lanes)
#define EXMI_ZPROD(NN, TT)                                                                \
  case NN:                                                                                \
    switch (a.kind) {                                                                     \
      case NL_GRADNORM: return go(zprod3d_kernel<NN, 8, NL_GRADNORM>, 256 / TT);           \
      case NL_GENERAL: return go(zprod3d_kernel<NN, 8, NL_GENERAL>, 256 / TT);             \
      case NL_POLY: return go(zprod3d_kernel<NN, 8, NL_POLY>, 256 / TT);                   \
      case NL_GRAY_SCOTT: return go(zprod3d_kernel<NN, 8, NL_GRAY_SCOTT>, 256 / TT);       \
      case NL_BZ: return go(zprod3d_kernel<NN, 8, NL_BZ>, 256 / TT);                       \
      case NL_CONV_CONS: return go(zprod3d_kernel<NN, 4, NL_CONV_CONS>, 256 / (2 * TT));   \
      case NL_CONV_NONCONS: return go(zprod3d_kernel<NN, 4, NL_CONV_NONCONS>, 256 / (2 * TT)); \
      default: return hipErrorInvalidValue;                                               \
    }
  switch (N) {
    EXMI_ZPROD(16, 1)
    EXMI_ZPROD(32, 2)
    EXMI_ZPROD(64, 4)
    EXMI_ZPROD(128, 8)
    default: return hipErrorInvalidValue;
  }
#undef EXMI_ZPROD
}

// The y lines of the nonlinear evaluation fused with its z rows for
// N <= 64: one workgroup owns the plane (trajectory, ix) after the inverse
// x lines, runs the NIN (kcut + 1) inverse y lines of the trajectory's
// fields on the planes kz <= kcut (times 1/N) into LDS, then per row iy the
// fields' c2r transforms along z, the products at every point (zpoint), the
// products' r2c transforms back into LDS, then the NOUT (kcut + 1) forward
// y lines into P, where the forward y line pass would leave them: the
// inverse y pass, the z-row kernel and the forward y pass in one. kcut <
// N/3 + 1 planes are held (the 2/3 rule's cutoff); the caller checks. P may
// equal F (a workgroup reads its whole plane before it writes it).
template <int N, int KIND>
struct YZCfg3 {
  using Z = ZProd<KIND>;
  static constexpr int NIN = Z::C * Z::NF, NOUT = Z::C * Z::NP;
  static constexpr int M = N / 2, NZ = M + 1, KCMAX = N / 3 + 1;
  static constexpr int PT1 = LineCfg<N>::PT, T1 = N / PT1, G1 = 256 / T1;  // y lines
  static constexpr int B1 = (NIN * KCMAX + G1 - 1) / G1;                    // inverse y lines per group
  static constexpr int B3 = (NOUT * KCMAX + G1 - 1) / G1;                   // forward y lines per group
  static constexpr int PT2 = 8, T2 = M / PT2, G2 = 256 / T2;                 // z rows
  static constexpr int L1 = G1 * LdsSize<N>::value, L2 = G2 * LdsSize<M>::value;
  static constexpr int LY = (NIN > NOUT ? NIN : NOUT) * N * KCMAX;
  static constexpr int LDS = LY > L1 ? (LY > L2 ? LY : L2) : (L1 > L2 ? L1 : L2);
  static_assert(N <= G2, "one z row per group");
};

template <int N, int KIND>
__global__ void __launch_bounds__(256) yzprod3d_kernel(NL3DArgs a, const cf* F, cf* P, int kcut,
                                                       const cf* __restrict__ tw, RealTables rt) {
  using C = YZCfg3<N, KIND>;
  constexpr int NIN = C::NIN, NOUT = C::NOUT;
  constexpr int M = C::M, NZ = C::NZ, KC = C::KCMAX, PT1 = C::PT1, T1 = C::T1, G1 = C::G1, B1 = C::B1;
  constexpr int PT2 = C::PT2, T2 = C::T2, B3 = C::B3;
  const long long Nr = a.Nr;
  // the y lines' scratch, the y-transformed fields Y[q][iy][kz], the z rows'
  // scratch, the z-transformed products, the y lines' scratch, in turn
  // (barriers between)
  __shared__ cf lds[C::LDS];
  const long long plane = blockIdx.x;  // traj * N + ix
  const long long traj = plane / N;
  const int ix = (int)(plane - traj * N);
  constexpr float rn = 1.f / (float)N;
  {
    const int nl = NIN * (kcut + 1);
    const int c = threadIdx.x / T1, j = threadIdx.x % T1;
    const Lane<N, PT1> ln(j);
    TwRegs<N, PT1> tr;
    tr.template load_pass<1>(j, tw);
    cf v[B1][PT1];
#pragma unroll
    for (int b = 0; b < B1; ++b) {
      const int l = c + b * G1;
      const int q = l < nl ? l / (kcut + 1) : 0, kz = l < nl ? l % (kcut + 1) : 0;
      const cf* x = F + (traj * NIN + q) * Nr + (long long)ix * N * NZ + kz;
#pragma unroll
      for (int s = 0; s < PT1; ++s) v[b][s] = l < nl ? x[(j + s * T1) * NZ] : cf{0.f, 0.f};
    }
    cfft<N, PT1, true, B1>(v, ln, lds + c * LdsSize<N>::value, tr);
    __syncthreads();  // every group is done with its scratch before Y overwrites it
#pragma unroll
    for (int b = 0; b < B1; ++b) {
      const int l = c + b * G1;
      if (l < nl) {
        const int q = l / (kcut + 1), kz = l % (kcut + 1);
#pragma unroll
        for (int s = 0; s < PT1; ++s) lds[(q * N + j + s * T1) * KC + kz] = rn * v[b][s];
      }
    }
  }
  __syncthreads();
  {
    const int g = threadIdx.x / T2, j = threadIdx.x % T2;
    const int iy = g < N ? g : 0;  // groups past the last row mirror row 0 and store nothing
    cf v[NIN][PT2];
    cf nyq[NIN];
#pragma unroll
    for (int q = 0; q < NIN; ++q) {
#pragma unroll
      for (int s = 0; s < PT2; ++s)
        v[q][s] = j + s * T2 <= kcut ? lds[(q * N + iy) * KC + j + s * T2] : cf{0.f, 0.f};
      nyq[q] = cf{0.f, 0.f};  // kcut < KC <= M
    }
    __syncthreads();  // every row is loaded before the z scratch overwrites Y
    const Lane<M, PT2> ln(j);
    TwRegs<M, PT2> tr;
    tr.load(j, rt);
    cf* scr = lds + g * LdsSize<M>::value;
    irfft<M, PT2, NIN>(v, nyq, ln, scr, tr);  // N times the physical pairs
    cf pr[NOUT][PT2];
#pragma unroll
    for (int s = 0; s < PT2; ++s) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // the two real points of the slot
        float f[NIN], o[NOUT];
#pragma unroll
        for (int q = 0; q < NIN; ++q) f[q] = (h ? v[q][s].im : v[q][s].re) * rn;
        zpoint<KIND>(a, f, o);
#pragma unroll
        for (int q = 0; q < NOUT; ++q) {
          if (h) pr[q][s].im = o[q];
          else pr[q][s].re = o[q];
        }
      }
    }
    cf pn[NOUT];
    rfft<M, PT2, NOUT>(pr, pn, ln, scr, tr, kcut);  // 2 X
    __syncthreads();  // every row is done with its scratch before Y2 overwrites it
    if (g < N) {
#pragma unroll
      for (int q = 0; q < NOUT; ++q) {
#pragma unroll
        for (int s = 0; s < PT2; ++s)
          if (j + s * T2 <= kcut) lds[(q * N + iy) * KC + j + s * T2] = 0.5f * pr[q][s];
      }
    }
  }
  __syncthreads();
  {
    // the forward y lines of the products' planes kz <= kcut
    const int nl3 = NOUT * (kcut + 1);
    const int c3 = threadIdx.x / T1, j3 = threadIdx.x % T1;
    cf w[B3][PT1];
#pragma unroll
    for (int b = 0; b < B3; ++b) {
      const int l = c3 + b * G1;
      const int q = l < nl3 ? l / (kcut + 1) : 0, kz = l < nl3 ? l % (kcut + 1) : 0;
#pragma unroll
      for (int s = 0; s < PT1; ++s) w[b][s] = l < nl3 ? lds[(q * N + j3 + s * T1) * KC + kz] : cf{0.f, 0.f};
    }
    __syncthreads();  // every line is loaded before the scratch overwrites Y2
    const Lane<N, PT1> ln3(j3);
    TwRegs<N, PT1> tr3;
    tr3.template load_pass<1>(j3, tw);
    cfft<N, PT1, false, B3>(w, ln3, lds + c3 * LdsSize<N>::value, tr3);
#pragma unroll
    for (int b = 0; b < B3; ++b) {
      const int l = c3 + b * G1;
      if (l < nl3) {
        const int q = l / (kcut + 1), kz = l % (kcut + 1);
        cf* y = P + (traj * NOUT + q) * Nr + (long long)ix * N * NZ + kz;
#pragma unroll
        for (int s = 0; s < PT1; ++s) y[(j3 + s * T1) * NZ] = w[b][s];
      }
    }
  }
}

// The fused y lines and z rows over the trajectories' N planes each (N = 16,
// 32 or 64 with kcut < N/3 + 1; the projected convection, gradient norm,
// general form, polynomial and coupled reactions); hipErrorInvalidValue
// otherwise (the caller runs the passes apart).
hipError_t launch_yzprod3d(const NL3DArgs& a, const cf* F, cf* P, int kcut, const cf* tw, const RealTables& rt,
                           hipStream_t st) {
  const int N = a.N;
  if (N > 64 || a.dims != 3 || kcut < 0 || kcut >= N / 3 + 1) return hipErrorInvalidValue;
  const int C = a.kind == NL_GRAY_SCOTT ? 2 : (a.kind == NL_BZ || a.kind == NL_PROJ3D) ? 3 : 1;
  const long long nplanes = a.ninst / C * N;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)nplanes), dim3(256), 0, st, a, F, P, kcut, tw, rt);
    return hipGetLastError();
  };
#define EXMI_YZ(NN)                                                         \
  case NN:                                                                  \
    switch (a.kind) {                                                       \
      case NL_PROJ3D: return go(yzprod3d_kernel<NN, NL_PROJ3D>);            \
      case NL_GRADNORM: return go(yzprod3d_kernel<NN, NL_GRADNORM>);        \
      case NL_GENERAL: return go(yzprod3d_kernel<NN, NL_GENERAL>);          \
      case NL_POLY: return go(yzprod3d_kernel<NN, NL_POLY>);                \
      case NL_GRAY_SCOTT: return go(yzprod3d_kernel<NN, NL_GRAY_SCOTT>);    \
      case NL_BZ: return go(yzprod3d_kernel<NN, NL_BZ>);                    \
      default: return hipErrorInvalidValue;                                 \
    }
  switch (N) {
    EXMI_YZ(16)
    EXMI_YZ(32)
    EXMI_YZ(64)
    default: return hipErrorInvalidValue;
  }
#undef EXMI_YZ
}

hipError_t launch_zcross3d(int N, const cf* F, cf* P, long long ntraj, long long Nr, int kcut, const RealTables& rt,
                           hipStream_t st) {
  auto go = [&](auto kern, int G) {
    const long long nrows = ntraj * N * N;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nrows + G - 1) / G)), dim3(256), 0, st, F, P, nrows, Nr, kcut, rt);
    return hipGetLastError();
  };
  switch (N) {
    case 16: return go(zcross3d_kernel<16, 8>, 256 / 1);
    case 32: return go(zcross3d_kernel<32, 8>, 256 / 2);
    case 64: return go(zcross3d_kernel<64, 8>, 256 / 4);
    case 128: return go(zcross3d_kernel<128, 8>, 256 / 8);
    case 256: return go(zcross3d_kernel<256, 8>, 256 / 16);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------ adjoint of the nonlinear term
// J(w)^T mu on whole spectra in three passes around the transforms, the
// transposes of the forward passes above under the inner products sum(f g)
// (fields) and sum(irfftn(a) irfftn(b)) (spectra), in which rfftn and
// irfftn are mutual adjoints, a Hermitian per-mode factor has adjoint its
// conjugate (i kappa -> -i kappa), and the dealias mask P and the mean
// removal Z are self-adjoint:
//   cofields : the cotangent fields of mu (P mu, conj-derivatives, Z)   -> H
//   (irfftn of every cotangent field, and of the forward fields of w)
//   products : the pointwise adjoint products of the two                -> Q
//   (rfftn of every product)
//   combine  : conj-derivative multiplies, scales, P                    -> out
// Per kind (h = irfftn(P mu); w, g_d = d_d w, u, v, omega_x, omega_y the
// forward fields of w):
//   gradient norm   -b/2 P Z F[sum g_d^2]      -> -b P sum_d conj(D_d) F[g_d irfftn(Z P mu)]
//   convection, conservative, C = D channels   -> -(b/2) P F[sum_j (h_aj + h_ja) w_j],
//                                                  h_cj = irfftn(P conj(D_j) mu_c)
//   convection, u . grad u                     -> -b P (F[sum_c h_c d_a w_c] + sum_j conj(D_j) F[h_a w_j])
//   polynomial                                 -> P F[p'(w) h]
//   Cahn-Hilliard s Lap P F[w^3]                 -> P F[3 w^2 irfftn(s Lap P mu)]
//   Gray-Scott, Belousov-Zhabotinsky           -> P F[sum_j dN_j/dw_c h_j] (the reactions'
//                                                  pointwise Jacobians, transposed)
//   general                                    -> P (F[w (2 g0 h + g1 hd)] + g2 sum_d conj(D_d) F[g_d hz]),
//                                                  hd = irfftn(P conj(sum_d D_d) mu), hz = irfftn(Z P mu);
//                                                  single-channel u sum_d d_d u: g0 (F[h sum_d g_d] + conj(sum_d D_d) F[h w])
//   vorticity (2-D) -b (u omega_x + v omega_y) -> -b P (conj(D_y) Linv F[omega_x h] - conj(D_x) Linv F[omega_y h]
//                                                        + conj(D_x) F[u h] + conj(D_y) F[v h])
//   projected convection Pl P F[u x omega],    -> P (F[omega x h] + i kappa x F[h x u]),
//   omega = irfftn(P i kappa x w_hat)             h = irfftn(Pl P mu) (the Leray projector Pl is
//                                                  real and symmetric per mode; the cross product
//                                                  with i kappa is its own conjugate transpose)
// (oracle/adjoint.py restates each from the forward formulas of
// nonlin_fun/_gradient_norm.py, _convection.py, _polynomial.py,
// _general_nonlinear.py, _vorticity_convection.py:78-99 and
// _projected_convection.py:19-136 with _leray.py:8-103.)
int adj3d_fields(int kind, int dims, int poly_lap, int* nh, int* nq) {
  (void)poly_lap;
  *nh = 0;
  *nq = 0;
  switch (kind) {
    case NL_GRADNORM: *nh = 1; *nq = dims; break;
    case NL_CONV_CONS: *nh = dims; *nq = 1; break;
    case NL_CONV_NONCONS: *nh = 1; *nq = 1 + dims; break;
    case NL_POLY:  // Cahn-Hilliard (poly_lap) included
    case NL_GRAY_SCOTT:
    case NL_BZ:
      *nh = 1;
      *nq = 1;
      break;
    case NL_GENERAL: *nh = 3; *nq = 2 + dims; break;
    case NL_VORT: *nh = 1; *nq = 4; break;
    case NL_PROJ3D: *nh = 1; *nq = 2; break;
    default: return 0;
  }
  return *nh;
}

__global__ void adj_cofields3d_kernel(NL3DArgs a, const cf* __restrict__ MU, cf* __restrict__ H) {
  const long long total = a.ninst * a.Nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / a.Nr, k = idx % a.Nr;
    const Mode3D m = mode3d(a, k);
    const cf mu = m.keep ? MU[idx] : cf{0.f, 0.f};
    cf* h = H + inst * a.nh * a.Nr + k;
    switch (a.kind) {
      case NL_GRADNORM: h[0] = (a.zero_mode_fix && m.zero) ? cf{0.f, 0.f} : mu; break;
      case NL_POLY:  // Cahn-Hilliard: the (real) scale times Laplacian first
        h[0] = a.poly_lap ? (a.scale * -(m.kx * m.kx + m.ky * m.ky + m.kz * m.kz)) * mu : mu;
        break;
      case NL_CONV_CONS:
        for (int j = 0; j < a.dims; ++j) h[j * a.Nr] = imul(-kap_d(a, m, j), mu);
        break;
      case NL_GENERAL: {
        h[0] = mu;
        h[a.Nr] = imul(-(m.kx + m.ky + m.kz), mu);
        h[2 * a.Nr] = (a.zero_mode_fix && m.zero) ? cf{0.f, 0.f} : mu;
        break;
      }
      case NL_PROJ3D: {
        // the Leray projection of the trajectory's three masked cotangents
        // (as in combine_at; the projector is its own transpose)
        const long long b = (inst - inst % a.channels) * a.Nr + k;
        const cf z{0.f, 0.f};
        const cf m0 = m.keep ? MU[b] : z, m1 = m.keep ? MU[b + a.Nr] : z, m2 = m.keep ? MU[b + 2 * a.Nr] : z;
        const cf div = imul(m.kx, m0) + imul(m.ky, m1) + imul(m.kz, m2);
        const float k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
        const cf pr = k2 != 0.f ? (1.f / k2) * div : z;
        const int c = (int)(inst % a.channels);
        h[0] = mu + imul(c == 0 ? m.kx : c == 1 ? m.ky : m.kz, pr);
        break;
      }
      default: h[0] = mu;
    }
  }
}

__global__ void adj_products3d_kernel(NL3DArgs a, const float* __restrict__ R, const float* __restrict__ HR,
                                      float* __restrict__ Q) {
  const long long total = a.ninst * a.N3;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / a.N3, p = idx % a.N3;
    const long long base = inst - inst % a.channels;
    const int ch = (int)(inst % a.channels);
    auto fld = [&](long long i, int f) { return R[(i * a.nf + f) * a.N3 + p]; };
    auto cot = [&](long long i, int f) { return HR[(i * a.nh + f) * a.N3 + p]; };
    float* q = Q + inst * a.nq * a.N3 + p;
    switch (a.kind) {
      case NL_GRADNORM: {
        const float h = cot(inst, 0);
        for (int d = 0; d < a.dims; ++d) q[(long long)d * a.N3] = fld(inst, d) * h;
        break;
      }
      case NL_CONV_CONS: {
        float acc = 0.f;
        for (int j = 0; j < a.dims; ++j) acc = fmaf(cot(inst, j) + cot(base + j, ch), fld(base + j, 0), acc);
        q[0] = acc;
        break;
      }
      case NL_CONV_NONCONS: {
        const float ha = cot(inst, 0);
        float acc = 0.f;
        for (int c = 0; c < a.dims; ++c) acc = fmaf(cot(base + c, 0), fld(base + c, 1 + ch), acc);
        q[0] = acc;
        for (int j = 0; j < a.dims; ++j) q[(long long)(1 + j) * a.N3] = ha * fld(base + j, 0);
        break;
      }
      case NL_POLY: {
        const float u = fld(inst, 0);
        float dp = 0.f, pw = 1.f;
        for (int i = 1; i < a.npoly; ++i) {
          dp = fmaf((float)i * a.poly[i], pw, dp);
          pw *= u;
        }
        q[0] = dp * cot(inst, 0);
        break;
      }
      case NL_GENERAL: {
        const float w = fld(inst, 0), h = cot(inst, 0), hd = cot(inst, 1), hz = cot(inst, 2);
        if (a.prod_nc) {
          float s = 0.f;
          for (int d = 0; d < a.dims; ++d) s += fld(inst, 1 + d);
          q[0] = a.g0 * h * s;
          q[a.N3] = a.g0 * h * w;
        } else {
          q[0] = w * fmaf(2.f * a.g0, h, a.g1 * hd);
          q[a.N3] = 0.f;
        }
        for (int d = 0; d < a.dims; ++d) q[(long long)(2 + d) * a.N3] = a.g2 * fld(inst, 1 + d) * hz;
        break;
      }
      case NL_VORT: {
        const float h = cot(inst, 0);
        q[0] = fld(inst, 1) * h;           // omega_x h
        q[a.N3] = fld(inst, 3) * h;        // omega_y h
        q[2 * a.N3] = fld(inst, 0) * h;    // u h
        q[3 * a.N3] = fld(inst, 2) * h;    // v h
        break;
      }
      case NL_GRAY_SCOTT: {
        // N0 = f (1 - u0) - u0 u1^2, N1 = -(f + k) u1 + u0 u1^2
        const float u0 = fld(base, 0), u1 = fld(base + 1, 0), h0 = cot(base, 0), h1 = cot(base + 1, 0);
        q[0] = ch == 0 ? fmaf(h1 - h0, u1 * u1, -a.feed * h0) : fmaf(2.f * u0 * u1, h1 - h0, -(a.feed + a.kill) * h1);
        break;
      }
      case NL_BZ: {
        // N0 = u0 + u1 - u0 u1 - u0^2, N1 = u2 - u1 - u0 u1, N2 = u0 - u2
        const float u0 = fld(base, 0), u1 = fld(base + 1, 0);
        const float h0 = cot(base, 0), h1 = cot(base + 1, 0), h2 = cot(base + 2, 0);
        q[0] = ch == 0 ? h0 * (1.f - u1 - 2.f * u0) - h1 * u1 + h2 : ch == 1 ? h0 * (1.f - u0) - h1 * (1.f + u0) : h1 - h2;
        break;
      }
      case NL_PROJ3D: {
        // fields u_c (0) and omega_c (1) of the trajectory's channels
        const long long c1 = base + (ch + 1) % 3, c2 = base + (ch + 2) % 3;
        q[0] = fld(c1, 1) * cot(c2, 0) - fld(c2, 1) * cot(c1, 0);     // (omega x h)_c
        q[a.N3] = cot(c1, 0) * fld(c2, 0) - cot(c2, 0) * fld(c1, 0);  // (h x u)_c
        break;
      }
      default: q[0] = 0.f;
    }
  }
}

__global__ void adj_combine3d_kernel(NL3DArgs a, const cf* __restrict__ QH, cf* __restrict__ out) {
  const long long total = a.ninst * a.Nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / a.Nr, k = idx % a.Nr;
    const Mode3D m = mode3d(a, k);
    const cf* q = QH + inst * a.nq * a.Nr + k;
    cf r = cf{0.f, 0.f};
    switch (a.kind) {
      case NL_GRADNORM:
        for (int d = 0; d < a.dims; ++d) r = r + imul(-kap_d(a, m, d), q[d * a.Nr]);
        r = -a.scale * r;
        break;
      case NL_CONV_CONS: r = (-0.5f * a.scale) * q[0]; break;
      case NL_CONV_NONCONS:
        r = q[0];
        for (int j = 0; j < a.dims; ++j) r = r + imul(-kap_d(a, m, j), q[(1 + j) * a.Nr]);
        r = -a.scale * r;
        break;
      case NL_GENERAL:
        r = q[0] + imul(-(m.kx + m.ky + m.kz), q[a.Nr]);
        for (int d = 0; d < a.dims; ++d) r = r + imul(-kap_d(a, m, d), q[(2 + d) * a.Nr]);
        break;
      case NL_VORT: {
        const float kx = m.ky, ky = m.kz;
        const float lap = -(kx * kx + ky * ky);
        const float inv = lap == 0.f ? 1.f : 1.f / lap;
        r = imul(-ky, inv * q[0]) + imul(kx, inv * q[a.Nr]) + imul(-kx, q[2 * a.Nr]) + imul(-ky, q[3 * a.Nr]);
        r = -a.scale * r;
        break;
      }
      case NL_PROJ3D: {
        // F[omega x h]_c + (i kappa x F[h x u])_c, the second from the
        // trajectory's other two channels
        const int c = (int)(inst % a.channels);
        const cf* q1 = QH + (inst - c) * a.nq * a.Nr + a.Nr + k;  // channel 0's F[h x u]
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const float k1 = c1 == 0 ? m.kx : c1 == 1 ? m.ky : m.kz, kk2 = c2 == 0 ? m.kx : c2 == 1 ? m.ky : m.kz;
        r = q[0] + imul(k1, q1[(long long)c2 * a.nq * a.Nr]) - imul(kk2, q1[(long long)c1 * a.nq * a.Nr]);
        break;
      }
      default: r = q[0];
    }
    out[idx] = m.keep ? r : cf{0.f, 0.f};
  }
}

__global__ void adj_linear3d_kernel(AdjLinArgs a) {
  const long long total = a.ninst * a.nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long inst = idx / a.nr, k = idx % a.nr;
    const long long traj = inst / a.channels, ch = inst % a.channels;
    const float* tb = a.tables + ((a.lin_batch > 1 ? traj : 0) * a.lin_channels + (a.lin_channels > 1 ? ch : 0)) *
                                     (long long)a.ncoef * a.nr;
    cf acc = cf{0.f, 0.f};
    for (int t = 0; t < a.nt; ++t) {
      const cf x = a.x[t][idx];
      const int op = a.op[t];
      cf v = x;
      if (op >= 100) {  // conj(E) x (100) or conj(E_half) x (101)
        const int ei_ = op - 100;
        const float er = tb[(long long)ei_ * a.nr + k];
        const float ei = a.e_im < 0 ? 0.f : tb[(long long)(a.e_im + ei_) * a.nr + k];
        v = cf{fmaf(er, x.re, ei * x.im), fmaf(er, x.im, -ei * x.re)};
      } else if (op >= 0) {  // a real coefficient of the tables
        v = tb[(long long)op * a.nr + k] * x;
      }
      acc = acc + a.mult[t] * v;
    }
    a.out[idx] = acc;
  }
}

static unsigned grid_of(long long total) {
  long long g = (total + 255) / 256;
  return (unsigned)(g > 65536 ? 65536 : g);
}

// The last two passes of irfftn for N <= 64, fused: one workgroup owns the
// plane (instance, ix) after the inverse x lines, runs its N/2 + 1 inverse
// y lines (times 1/N) into LDS, then its N c2r rows along z (times 1/N)
// from LDS into x; row t = plane * N + iy of the output starts at
// (t / rpg) gstride + (t % rpg) N (rpg > 0: a strided batch such as a
// trajectory buffer) or t N. Saves the y pass's store and the z pass's
// load of the plane.
template <int N>
struct YZCfg {
  static constexpr int M = N / 2, NZ = M + 1;
  static constexpr int PT1 = LineCfg<N>::PT, T1 = N / PT1, CPB1 = 256 / T1;  // y lines
  static constexpr int PT2 = 8, T2 = M / PT2, G2 = 256 / T2;                    // z rows
  static constexpr int L1 = CPB1 * LdsSize<N>::value, L2 = G2 * LdsSize<M>::value;
  static constexpr int SCR = L1 > L2 ? L1 : L2;
  static_assert(NZ <= CPB1 && N <= G2, "one y line per lane group and one z row per group");
};

template <int N>
__global__ void __launch_bounds__(256) yz_inverse3d_kernel(const cf* __restrict__ X, float* __restrict__ x,
                                                           long long nplanes, const cf* __restrict__ tw,
                                                           RealTables rt, long long rpg, long long gstride) {
  using C = YZCfg<N>;
  constexpr int M = C::M, NZ = C::NZ, PT1 = C::PT1, T1 = C::T1, PT2 = C::PT2, T2 = C::T2;
  __shared__ cf scr[C::SCR];
  __shared__ cf Y[N * NZ];
  const long long plane = blockIdx.x;
  const cf* src = X + plane * N * NZ;
  constexpr float rn = 1.f / (float)N;
  {
    // the inverse y line kz = c of the plane
    const int c = threadIdx.x / T1, j = threadIdx.x % T1;
    const int kz = c < NZ ? c : 0;  // groups past the last line mirror line 0 and store nothing
    const Lane<N, PT1> ln(j);
    TwRegs<N, PT1> tr;
    tr.template load_pass<1>(j, tw);
    cf v[1][PT1];
#pragma unroll
    for (int s = 0; s < PT1; ++s) v[0][s] = src[(j + s * T1) * NZ + kz];
    cfft<N, PT1, true, 1>(v, ln, scr + c * LdsSize<N>::value, tr);
    if (c < NZ) {
#pragma unroll
      for (int s = 0; s < PT1; ++s) Y[(j + s * T1) * NZ + kz] = rn * v[0][s];
    }
  }
  __syncthreads();
  {
    // the c2r row iy = g along z
    const int g = threadIdx.x / T2, j = threadIdx.x % T2;
    const int iy = g < N ? g : 0;  // groups past the last row mirror row 0 and store nothing
    const Lane<M, PT2> ln(j);
    TwRegs<M, PT2> tr;
    tr.load(j, rt);
    cf v[1][PT2];
    cf nyq[1];
#pragma unroll
    for (int s = 0; s < PT2; ++s) v[0][s] = Y[iy * NZ + j + s * T2];
    nyq[0] = Y[iy * NZ + M];
    irfft<M, PT2, 1>(v, nyq, ln, scr + g * LdsSize<M>::value, tr);  // N x
    if (g < N) {
      const long long t = plane * N + iy;
      const long long ro = rpg > 0 ? (t / rpg) * gstride + (t % rpg) * N : t * N;
      float2* y = reinterpret_cast<float2*>(x + ro);
#pragma unroll
      for (int s = 0; s < PT2; ++s) y[j + s * T2] = make_float2(v[0][s].re * rn, v[0][s].im * rn);
    }
  }
}

// The fused y and z passes of irfftn over nplanes = I N planes (N = 16, 32
// or 64); hipErrorInvalidValue for the other sizes.
hipError_t launch_yz_inverse3d(int N, const cf* X, float* x, long long nplanes, const cf* tw, const RealTables& rt,
                               long long rpg, long long gstride, hipStream_t st) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)nplanes), dim3(256), 0, st, X, x, nplanes, tw, rt, rpg, gstride);
    return hipGetLastError();
  };
  switch (N) {
    case 16: return go(yz_inverse3d_kernel<16>);
    case 32: return go(yz_inverse3d_kernel<32>);
    case 64: return go(yz_inverse3d_kernel<64>);
    default: return hipErrorInvalidValue;
  }
}

// The fields pass fused with the inverse x lines, one lane group per unit
// line: a unit is a trajectory for the projected convection (its three
// channels' spectra in, the six fields u_c, omega_c out) and an instance for
// the kinds whose fields are derivatives of the one spectrum (gradient
// norm: d_d u; non-conservative convection and the general form: u and
// d_d u). Each spectrum value is loaded once for all the unit's fields
// (fields_lines_kernel loads it once per field); the NOUT lines are then
// transformed together. Lines with iz >= kkeep do not run.
template <int KIND>
struct FieldUnit {
  static constexpr int CIN = KIND == NL_PROJ3D ? 3 : 1;  // spectra read per point
  static constexpr int NOUT = KIND == NL_PROJ3D ? 6 : KIND == NL_GRADNORM ? 3 : 4;
};

template <int N, int KIND>
__global__ void __launch_bounds__(256) fields_unit_lines_kernel(NL3DArgs a, const cf* __restrict__ S,
                                                                cf* __restrict__ F, float scale,
                                                                const cf* __restrict__ tw, int kkeep) {
  using C = XStageCfg<N>;
  using U = FieldUnit<KIND>;
  constexpr int PT = C::PT, T = C::T, CIN = U::CIN, NOUT = U::NOUT;
  constexpr int LDSN = LdsSize<N>::value;
  __shared__ cf lds_x[C::CPB * LDSN];
  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const int NZ = N / 2 + 1;
  const long long inner = (long long)N * NZ;
  const long long nrun = a.ninst / CIN * N * kkeep;
  const long long raw = (long long)blockIdx.x * C::CPB + c;
  const bool active = raw < nrun;
  const long long run = active ? raw : 0;  // inactive groups mirror line 0 and store nothing
  const long long unit = run / ((long long)N * kkeep);
  const int rr = (int)(run - unit * N * kkeep);
  const int iy = rr / kkeep, iz = rr - iy * kkeep;
  const long long li = (long long)iy * NZ + iz;
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, tw);
  cf v[NOUT][PT];
  const cf* src = S + unit * CIN * a.Nr + li;
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int ix = j + s * T;
    const Mode3D m = mode3d_at(a, ix, iy, iz);
    const cf z{0.f, 0.f};
    if constexpr (KIND == NL_PROJ3D) {
      const cf w0 = m.keep ? src[(long long)ix * inner] : z, w1 = m.keep ? src[a.Nr + (long long)ix * inner] : z,
               w2 = m.keep ? src[2 * a.Nr + (long long)ix * inner] : z;
      // field 2c + f of channel c: u_c, omega_c = (i kappa x u_hat)_c
      v[0][s] = w0;
      v[1][s] = imul(m.ky, w2) - imul(m.kz, w1);
      v[2][s] = w1;
      v[3][s] = imul(m.kz, w0) - imul(m.kx, w2);
      v[4][s] = w2;
      v[5][s] = imul(m.kx, w1) - imul(m.ky, w0);
    } else {
      const cf w = m.keep ? src[(long long)ix * inner] : z;
      if constexpr (KIND == NL_GRADNORM) {
        v[0][s] = imul(m.kx, w);
        v[1][s] = imul(m.ky, w);
        v[2][s] = imul(m.kz, w);
      } else {  // u and its three derivatives
        v[0][s] = w;
        v[1][s] = imul(m.kx, w);
        v[2][s] = imul(m.ky, w);
        v[3][s] = imul(m.kz, w);
      }
    }
  }
  cfft<N, PT, true, NOUT>(v, ln, lds, tr);
  // the stores run transposed through LDS: thread (q0, cl) stores the
  // points ix = q0 + q T of the workgroup's line cl, so that a wavefront
  // writes neighbouring lines (runs of up to kkeep entries) per point
  const int cl = threadIdx.x % C::CPB, q0 = threadIdx.x / C::CPB;
  const long long sraw = (long long)blockIdx.x * C::CPB + cl;
  const bool sact = sraw < nrun;
  const long long srun = sact ? sraw : 0;
  const long long sunit = srun / ((long long)N * kkeep);
  const int srr = (int)(srun - sunit * N * kkeep);
  const int siy = srr / kkeep, siz = srr - siy * kkeep;
  const long long sli = (long long)siy * NZ + siz;
#pragma unroll
  for (int q = 0; q < NOUT; ++q) {
    if (q > 0) __syncthreads();  // the previous field's reads are done
#pragma unroll
    for (int s = 0; s < PT; ++s) lds[j + s * T] = scale * v[q][s];
    __syncthreads();
    if (sact) {
      cf* o = F + (sunit * NOUT + q) * a.Nr + sli;
#pragma unroll
      for (int qq = 0; qq < PT; ++qq) {
        const int ix = q0 + qq * T;
        o[(long long)ix * inner] = lds_x[cl * LDSN + ix];
      }
    }
  }
}

// The unit-line fields pass for the kinds it covers (3-D, N <= 128);
// hipErrorInvalidValue for the others.
hipError_t launch_fields_unit_lines3d(const NL3DArgs& a, const cf* S, cf* F, float scale, const cf* tw, int kkeep,
                                      hipStream_t st) {
  if (a.dims != 3) return hipErrorInvalidValue;
  auto go = [&](auto kern, int cin, int cpb) {
    const long long nrun = a.ninst / cin * (long long)a.N * kkeep;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nrun + cpb - 1) / cpb)), dim3(256), 0, st, a, S, F, scale, tw, kkeep);
    return hipGetLastError();
  };
#define EXMI_FUNIT(NN)                                                                                  \
  case NN:                                                                                              \
    switch (a.kind) {                                                                                   \
      case NL_PROJ3D: return go(fields_unit_lines_kernel<NN, NL_PROJ3D>, 3, XStageCfg<NN>::CPB);         \
      case NL_GRADNORM: return go(fields_unit_lines_kernel<NN, NL_GRADNORM>, 1, XStageCfg<NN>::CPB);     \
      case NL_GENERAL: return go(fields_unit_lines_kernel<NN, NL_GENERAL>, 1, XStageCfg<NN>::CPB);       \
      case NL_CONV_NONCONS: return go(fields_unit_lines_kernel<NN, NL_CONV_NONCONS>, 1, XStageCfg<NN>::CPB); \
      default: return hipErrorInvalidValue;                                                             \
    }
  switch (a.N) {
    EXMI_FUNIT(16)
    EXMI_FUNIT(32)
    EXMI_FUNIT(64)
    EXMI_FUNIT(128)
    default: return hipErrorInvalidValue;
  }
#undef EXMI_FUNIT
}

hipError_t launch_fields_lines3d(const NL3DArgs& a, const cf* S, cf* F, float scale, const cf* tw, int kkeep,
                                 hipStream_t st) {
  auto go = [&](auto kern, int cpb) {
    const long long nrun = a.ninst * a.nf * (long long)a.N * kkeep;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nrun + cpb - 1) / cpb)), dim3(256), 0, st, a, S, F, scale, tw, kkeep);
    return hipGetLastError();
  };
  switch (a.N) {
    case 16: return go(fields_lines_kernel<16>, LineCfg<16>::CPB);
    case 32: return go(fields_lines_kernel<32>, LineCfg<32>::CPB);
    case 64: return go(fields_lines_kernel<64>, LineCfg<64>::CPB);
    case 128: return go(fields_lines_kernel<128>, LineCfg<128>::CPB);
    case 256: return go(fields_lines_kernel<256>, LineCfg<256>::CPB);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fields3d(const NL3DArgs& a, const cf* S, cf* F, hipStream_t st) {
  hipLaunchKernelGGL(fields3d_kernel, dim3(grid_of(a.ninst * a.Nr)), dim3(256), 0, st, a, S, F);
  return hipGetLastError();
}
hipError_t launch_products3d(const NL3DArgs& a, const float* R, float* PR, hipStream_t st) {
  hipLaunchKernelGGL(products3d_kernel, dim3(grid_of(a.ninst * a.N3)), dim3(256), 0, st, a, R, PR);
  return hipGetLastError();
}
hipError_t launch_adj_cofields3d(const NL3DArgs& a, const cf* MU, cf* H, hipStream_t st) {
  hipLaunchKernelGGL(adj_cofields3d_kernel, dim3(grid_of(a.ninst * a.Nr)), dim3(256), 0, st, a, MU, H);
  return hipGetLastError();
}
hipError_t launch_adj_products3d(const NL3DArgs& a, const float* R, const float* HR, float* Q, hipStream_t st) {
  hipLaunchKernelGGL(adj_products3d_kernel, dim3(grid_of(a.ninst * a.N3)), dim3(256), 0, st, a, R, HR, Q);
  return hipGetLastError();
}
hipError_t launch_adj_combine3d(const NL3DArgs& a, const cf* QH, cf* out, hipStream_t st) {
  hipLaunchKernelGGL(adj_combine3d_kernel, dim3(grid_of(a.ninst * a.Nr)), dim3(256), 0, st, a, QH, out);
  return hipGetLastError();
}
hipError_t launch_adj_linear3d(const AdjLinArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(adj_linear3d_kernel, dim3(grid_of(a.ninst * a.nr)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_combine3d(const NL3DArgs& a, const cf* PH, cf* Nh, hipStream_t st) {
  hipLaunchKernelGGL(combine3d_kernel, dim3(grid_of(a.ninst * a.Nr)), dim3(256), 0, st, a, PH, Nh);
  return hipGetLastError();
}

// ------------------------------------------------------ 3-D ETDRK stages
// The stage combinations of etdrk/_etdrk_{1,2,3,4}.py on whole spectra,
// coefficient c_i of the plan's tables (orders 1-2: [E, c1, c2] with Im E
// at e_im for a complex L; order 3: 7 entries, order 4: 6, as coef.hip
// forms them). X is the new state's spectrum.
// Stages that read the first nonlinear term N0, and those that read the
// latest one NC.
__device__ __forceinline__ bool stage_reads_n0(int m) {
  return m == ST1 || m == ST2A || m == ST2B || m == ST3A || m == ST3B || m == ST4A || m == ST4C;
}
__device__ __forceinline__ bool stage_reads_nc(int m) { return !(m == ST1 || m == ST2A || m == ST3A || m == ST4A); }

// The coefficient table of instance (trajectory traj, channel ch).
__device__ __forceinline__ const float* stage_table(const Stage3DArgs& a, long long traj, long long ch) {
  return a.tables +
         ((a.lin_batch > 1 ? traj : 0) * a.lin_channels + (a.lin_channels > 1 ? ch : 0)) * (long long)a.ncoef * a.nr;
}

// The stage combination of stage MODE at entry idx (mode k of its instance,
// tb the instance's coefficient table) with the nonlinear terms n0 (N0) and
// nc (NC) given; U0, S, S1, A as in Stage3DArgs. Returns the new S[idx].
template <int MODE>
__device__ __forceinline__ cf stage_core(int e_im, long long nr, const float* tb, long long k, long long idx,
                                         const cf* U0, cf* S, cf* S1, cf* A, cf n0, cf nc) {
  auto c = [&](int i) { return tb[(long long)i * nr + k]; };
  // coefficient i (0: E, 1: E_half for orders 3-4) times x, complex when L is
  auto cmul = [&](int i, cf x) -> cf {
    const float er = c(i);
    if (e_im < 0) return er * x;
    const float ei = c(e_im + i);
    return cf{fmaf(er, x.re, -ei * x.im), fmaf(er, x.im, ei * x.re)};
  };
  auto emul = [&](cf x) -> cf { return cmul(0, x); };
  cf r;
  if constexpr (MODE == ST1 || MODE == ST2A) {
    r = emul(U0[idx]) + c(1) * n0;
  } else if constexpr (MODE == ST2B) {
    r = S[idx] + c(2) * (nc - n0);
  } else if constexpr (MODE == ST3A) {
    const cf u0 = U0[idx];
    A[idx] = emul(u0) + c(4) * n0;
    r = cmul(1, u0) + c(2) * n0;
  } else if constexpr (MODE == ST3B) {
    const cf u0 = U0[idx];
    A[idx] = A[idx] + c(5) * nc;
    r = emul(u0) + c(3) * (2.f * nc - n0);
  } else if constexpr (MODE == ST3C) {
    r = A[idx] + c(6) * nc;
  } else if constexpr (MODE == ST4A) {
    const cf u0 = U0[idx];
    A[idx] = emul(u0) + c(3) * n0;
    r = cmul(1, u0) + c(2) * n0;
    S1[idx] = r;
  } else if constexpr (MODE == ST4B) {
    const cf u0 = U0[idx];
    A[idx] = A[idx] + (2.f * c(4)) * nc;
    r = cmul(1, u0) + c(2) * nc;
  } else if constexpr (MODE == ST4C) {
    A[idx] = A[idx] + (2.f * c(4)) * nc;
    r = cmul(1, S1[idx]) + c(2) * (2.f * nc - n0);
  } else {
    static_assert(MODE == ST4D, "stage mode");
    r = A[idx] + c(5) * nc;
  }
  S[idx] = r;
  return r;
}

// Calls f(std::integral_constant<int, mode>) for the runtime stage mode.
template <class F>
__device__ __forceinline__ void with_stage_mode(int mode, F&& f) {
  switch (mode) {
    case ST1: f(std::integral_constant<int, ST1>{}); break;
    case ST2A: f(std::integral_constant<int, ST2A>{}); break;
    case ST2B: f(std::integral_constant<int, ST2B>{}); break;
    case ST3A: f(std::integral_constant<int, ST3A>{}); break;
    case ST3B: f(std::integral_constant<int, ST3B>{}); break;
    case ST3C: f(std::integral_constant<int, ST3C>{}); break;
    case ST4A: f(std::integral_constant<int, ST4A>{}); break;
    case ST4B: f(std::integral_constant<int, ST4B>{}); break;
    case ST4C: f(std::integral_constant<int, ST4C>{}); break;
    case ST4D: f(std::integral_constant<int, ST4D>{}); break;
  }
}

// The stage combination at entry idx with the nonlinear terms n0 (N0) and
// nc (NC) given.
__device__ __forceinline__ void stage_at(const Stage3DArgs& a, long long idx, cf n0, cf nc) {
  const long long inst = idx / a.nr, k = idx % a.nr;
  const float* tb = stage_table(a, inst / a.channels, inst % a.channels);
  with_stage_mode(a.mode, [&](auto m) {
    stage_core<decltype(m)::value>(a.e_im, a.nr, tb, k, idx, a.U0, a.S, a.S1, a.A, n0, nc);
  });
}

__global__ void stage3d_kernel(Stage3DArgs a) {
  const long long total = a.ninst * a.nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const cf n0 = stage_reads_n0(a.mode) ? a.N0[idx] : cf{0.f, 0.f};
    const cf nc = stage_reads_nc(a.mode) ? a.NC[idx] : cf{0.f, 0.f};
    stage_at(a, idx, n0, nc);
  }
}

// The combine pass fused with the stage that consumes its result: N_hat is
// formed in registers from the product spectra PH and enters the stage as
// N0 (first: also stored to N0out for the later stages) or as NC.
__global__ void stage_nl3d_kernel(NL3DArgs na, Stage3DArgs a, const cf* __restrict__ PH, cf* __restrict__ N0out) {
  const long long total = a.ninst * a.nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const cf r = combine_at(na, PH, idx);
    if (N0out) {
      N0out[idx] = r;
      stage_at(a, idx, r, cf{0.f, 0.f});
    } else {
      stage_at(a, idx, stage_reads_n0(a.mode) ? a.N0[idx] : cf{0.f, 0.f}, r);
    }
  }
}

hipError_t launch_stage3d(const Stage3DArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(stage3d_kernel, dim3(grid_of(a.ninst * a.nr)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// The same for the projected convection (three channels): one thread per
// (trajectory, mode) projects the three product spectra once and runs the
// three channels' stages.
__global__ void stage_nl3d_proj_kernel(NL3DArgs na, Stage3DArgs a, const cf* __restrict__ PH, cf* __restrict__ N0out) {
  const long long total = a.ninst / 3 * a.nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long traj = idx / a.nr, k = idx % a.nr;
    const Mode3D m = mode3d(na, k);
    const long long b = traj * 3 * a.nr + k;
    const cf p0 = PH[b], p1 = PH[b + a.nr], p2 = PH[b + 2 * a.nr];
    // Leray projection (_leray.py:86-103): c + i kappa_c p with
    // p = -div / Laplace = (i kappa . c) / |kappa|^2 (p = 0 at the mean mode)
    const cf div = imul(m.kx, p0) + imul(m.ky, p1) + imul(m.kz, p2);
    const float k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
    const cf pr = k2 != 0.f ? (1.f / k2) * div : cf{0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const cf pc = c == 0 ? p0 : c == 1 ? p1 : p2;
      cf r = m.keep ? pc + imul(c == 0 ? m.kx : c == 1 ? m.ky : m.kz, pr) : cf{0.f, 0.f};
      if (c == 0 && m.inj) r = r + cf{na.inj_value, 0.f};
      const long long i = b + c * a.nr;
      if (N0out) {
        N0out[i] = r;
        stage_at(a, i, r, cf{0.f, 0.f});
      } else {
        stage_at(a, i, stage_reads_n0(a.mode) ? a.N0[i] : cf{0.f, 0.f}, r);
      }
    }
  }
}

// The forward x lines of the projected convection's three product spectra
// fused with the projection and the stage: one lane group owns the line
// (trajectory, iy, kz) of all three channels (P holds them after the z rows
// and the forward y lines), transforms them along x (planes kz > kcut hold
// no products and are not read: their N_hat is zero), and applies at every
// point the projection and the stage of stage_nl3d_proj_kernel. The state
// buffers come in as separate non-aliasing arguments so that the stage's
// loads of later points need not wait for the stores of earlier ones.

template <int N, bool EMIT>
__global__ void __launch_bounds__(256)
    xstage_proj3d_kernel(NL3DArgs na, Stage3DArgs a, const cf* __restrict__ P, cf* __restrict__ N0out, int kcut,
                         const cf* __restrict__ tw, const cf* __restrict__ U0, const cf* __restrict__ N0,
                         cf* __restrict__ S, cf* __restrict__ S1, cf* __restrict__ A, cf* __restrict__ K) {
  using C = XStageCfg<N>;
  constexpr int PT = C::PT, T = C::T;
  constexpr int LDSN = LdsSize<N>::value;
  __shared__ cf lds_x[C::CPB * LDSN];
  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const int NZ = N / 2 + 1;
  const long long inner = (long long)N * NZ;  // lines (iy, kz) per trajectory
  const long long nlines = a.ninst / 3 * inner;
  const long long raw = (long long)blockIdx.x * C::CPB + c;
  const long long line = raw < nlines ? raw : 0;  // inactive groups mirror line 0 and store nothing
  const long long traj = line / inner;
  const int li = (int)(line - traj * inner);
  const int iy = li / NZ, iz = li - iy * NZ;
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, tw);
  cf v[3][PT];
  const long long b0 = traj * 3 * a.nr + li;  // channel 0's entry of this line at ix = 0
  // The stage (and, in the emitting variant, the loads; the non-emitting
  // variant measured 1.4 us slower with them) runs transposed through LDS: thread
  // (q0, cl) takes the points ix = q0 + q T of the workgroup's line cl, so
  // that a wavefront's accesses cover neighbouring lines (contiguous
  // entries).
  const int cl = threadIdx.x % C::CPB, q0 = threadIdx.x / C::CPB;
  const long long sraw = (long long)blockIdx.x * C::CPB + cl;
  const bool sact = sraw < nlines;
  const long long sline = sact ? sraw : 0;
  const long long straj = sline / inner, sli = sline - straj * inner;
  const long long sb0 = straj * 3 * a.nr + sli;
  const bool skept = sli % NZ <= kcut;
  if constexpr (EMIT) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      if (ch > 0) __syncthreads();  // the previous channel's reads are done
#pragma unroll
      for (int q = 0; q < PT; ++q) {
        const int ix = q0 + q * T;
        lds_x[cl * LDSN + ix] = skept ? P[sb0 + ch * a.nr + (long long)ix * inner] : cf{0.f, 0.f};
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < PT; ++s) v[ch][s] = lds[j + s * T];
    }
  } else {
    const bool kept = iz <= kcut;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
      for (int s = 0; s < PT; ++s)
        v[ch][s] = kept ? P[b0 + ch * a.nr + (long long)(j + s * T) * inner] : cf{0.f, 0.f};
    }
  }
  cfft<N, PT, false, 3>(v, ln, lds, tr);  // only this group uses its LDS line until the next barrier
  // the projection (as in stage_nl3d_proj_kernel), in place
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const Mode3D m = mode3d_at(na, j + s * T, iy, iz);
    const cf p0 = v[0][s], p1 = v[1][s], p2 = v[2][s];
    const cf div = imul(m.kx, p0) + imul(m.ky, p1) + imul(m.kz, p2);
    const float k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
    const cf pr = k2 != 0.f ? (1.f / k2) * div : cf{0.f, 0.f};
    const cf z{0.f, 0.f};
    v[0][s] = m.keep ? p0 + imul(m.kx, pr) : z;
    v[1][s] = m.keep ? p1 + imul(m.ky, pr) : z;
    v[2][s] = m.keep ? p2 + imul(m.kz, pr) : z;
    if (m.inj) v[0][s] = v[0][s] + cf{na.inj_value, 0.f};
  }
  // one copy of the stage loop per stage mode (a per-point switch would put
  // every mode's code into every point's path)
  // With EMIT (the step's last stage of a rollout that emits this state)
  // the new state's values go back through LDS into the lane groups' line
  // order, and their inverse x lines (times 1/N, every kz plane: the first
  // pass of irfftn) are stored to K.
  cf e[EMIT ? 3 : 1][PT];
  with_stage_mode(a.mode, [&](auto mc) {
    constexpr int M = decltype(mc)::value;
    constexpr bool first = !(M == ST2B || M == ST3B || M == ST3C || M == ST4B || M == ST4C || M == ST4D);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      if (ch > 0) __syncthreads();  // the previous channel's reads are done
#pragma unroll
      for (int s = 0; s < PT; ++s) lds[j + s * T] = v[ch][s];
      __syncthreads();
      const float* tb = stage_table(a, straj, ch);
#pragma unroll
      for (int q = 0; q < PT; ++q) {
        const int ix = q0 + q * T;
        const cf r = lds_x[cl * LDSN + ix];
        const long long k = sli + (long long)ix * inner;
        const long long i = sb0 + ch * a.nr + (long long)ix * inner;
        if (!sact) continue;
        cf sv;
        if constexpr (first) {
          N0out[i] = r;
          sv = stage_core<M>(a.e_im, a.nr, tb, k, i, U0, S, S1, A, r, cf{0.f, 0.f});
        } else {
          sv = stage_core<M>(a.e_im, a.nr, tb, k, i, U0, S, S1, A, stage_reads_n0(M) ? N0[i] : cf{0.f, 0.f}, r);
        }
        if constexpr (EMIT) lds_x[cl * LDSN + ix] = sv;  // the slot this thread alone read
      }
      if constexpr (EMIT) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < PT; ++s) e[ch][s] = lds[j + s * T];
      }
    }
  });
  if constexpr (EMIT) {
    cfft<N, PT, true, 3>(e, ln, lds, tr);  // only this group uses its LDS line until the next barrier
    // stored transposed through LDS, as the loads
    constexpr float rn = 1.f / (float)N;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      if (ch > 0) __syncthreads();  // the previous channel's reads are done
#pragma unroll
      for (int s = 0; s < PT; ++s) lds[j + s * T] = rn * e[ch][s];
      __syncthreads();
      if (sact) {
#pragma unroll
        for (int q = 0; q < PT; ++q) {
          const int ix = q0 + q * T;
          K[sb0 + ch * a.nr + (long long)ix * inner] = lds_x[cl * LDSN + ix];
        }
      }
    }
  }
}

// K: where the new state's inverse x lines go (the emitting last stage of a
// rollout step), or nullptr.
hipError_t launch_xstage_proj3d(const NL3DArgs& na, const Stage3DArgs& a, const cf* P, cf* N0out, int kcut,
                                const cf* tw, hipStream_t st, cf* K) {
  const bool first = a.mode == ST1 || a.mode == ST2A || a.mode == ST3A || a.mode == ST4A;
  if (first != (N0out != nullptr)) return hipErrorInvalidValue;  // the first stage keeps N0, the others read it
  auto go = [&](auto kern, int cpb) {
    const long long nlines = a.ninst / 3 * (long long)na.N * na.NZ;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nlines + cpb - 1) / cpb)), dim3(256), 0, st, na, a, P, N0out, kcut,
                       tw, a.U0, a.N0, a.S, a.S1, a.A, K);
    return hipGetLastError();
  };
#define EXMI_XSP(NN)                                                                                 \
  case NN:                                                                                           \
    return K ? go(xstage_proj3d_kernel<NN, true>, XStageCfg<NN>::CPB)                                \
             : go(xstage_proj3d_kernel<NN, false>, XStageCfg<NN>::CPB);
  switch (na.N) {
    EXMI_XSP(16)
    EXMI_XSP(32)
    EXMI_XSP(64)
    EXMI_XSP(128)
    default: return hipErrorInvalidValue;
  }
#undef EXMI_XSP
}

// N_hat at one mode from the instance's product spectra values ph[0..NP) of
// the kinds with fused z rows: combine_at's formulas for these kinds, the
// kind fixed at compile time.
template <int KIND>
__device__ __forceinline__ cf combine_vals(const NL3DArgs& a, const Mode3D& m, const cf* ph) {
  const cf ph0 = ph[0];
  cf r;
  if constexpr (KIND == NL_CONV_CONS) {
    r = (-0.5f * a.scale) * (imul(m.kx, ph[0]) + imul(m.ky, ph[1]) + imul(m.kz, ph[2]));
  } else if constexpr (KIND == NL_CONV_NONCONS) {
    r = -a.scale * ph0;
  } else if constexpr (KIND == NL_POLY) {
    r = a.poly_lap ? (a.scale * -(m.kx * m.kx + m.ky * m.ky + m.kz * m.kz)) * ph0 : ph0;
  } else if constexpr (KIND == NL_GRADNORM) {
    r = (a.zero_mode_fix && m.zero) ? cf{0.f, 0.f} : (-0.5f * a.scale) * ph0;
  } else if constexpr (KIND == NL_GENERAL) {
    const float cg = (a.zero_mode_fix && m.zero) ? 0.f : 0.5f * a.g2;
    r = a.g0 * ph0 + imul(0.5f * a.g1 * (m.kx + m.ky + m.kz), ph0) + cg * ph[1];
  } else {
    r = ph0;  // the coupled reactions
  }
  return m.keep ? r : cf{0.f, 0.f};
}

// The forward x lines of the product spectra of the kinds with fused z rows
// (gradient norm, polynomial, general form, coupled reactions, three-channel
// convections) fused with
// the combine and the stage, as xstage_proj3d_kernel does for the projected
// convection: a lane group owns the line (instance, iy, kz) of the
// instance's NP product spectra, transforms them along x (planes kz > kcut
// are not read), forms N_hat per point, and the workgroup runs the stage
// transposed through LDS, one loop per stage mode.
template <int N, int KIND, bool EMIT>
__global__ void __launch_bounds__(256)
    xstage3d_kernel(NL3DArgs na, Stage3DArgs a, const cf* __restrict__ P, cf* __restrict__ N0out, int kcut,
                    const cf* __restrict__ tw, const cf* __restrict__ U0, const cf* __restrict__ N0,
                    cf* __restrict__ S, cf* __restrict__ S1, cf* __restrict__ A, cf* __restrict__ K) {
  using C = XStageCfg<N>;
  constexpr int PT = C::PT, T = C::T, NP = ZProd<KIND>::NP;
  constexpr int LDSN = LdsSize<N>::value;
  __shared__ cf lds_x[C::CPB * LDSN];
  const int c = threadIdx.x / T, j = threadIdx.x % T;
  const int NZ = N / 2 + 1;
  const long long inner = (long long)N * NZ;  // lines (iy, kz) per instance
  const long long nlines = a.ninst * inner;
  const long long raw = (long long)blockIdx.x * C::CPB + c;
  const long long line = raw < nlines ? raw : 0;  // inactive groups mirror line 0 and store nothing
  const long long inst = line / inner;
  const int li = (int)(line - inst * inner);
  const int iy = li / NZ, iz = li - iy * NZ;
  const bool kept = iz <= kcut;
  cf* lds = lds_x + c * LDSN;
  const Lane<N, PT> ln(j);
  TwRegs<N, PT> tr;
  tr.template load_pass<1>(j, tw);
  cf v[NP][PT];
  const long long b0 = inst * NP * a.nr + li;  // product 0 of this line at ix = 0
#pragma unroll
  for (int q = 0; q < NP; ++q) {
#pragma unroll
    for (int s = 0; s < PT; ++s)
      v[q][s] = kept ? P[b0 + q * a.nr + (long long)(j + s * T) * inner] : cf{0.f, 0.f};
  }
  cfft<N, PT, false, NP>(v, ln, lds, tr);
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    cf ph[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) ph[q] = v[q][s];
    v[0][s] = combine_vals<KIND>(na, mode3d_at(na, j + s * T, iy, iz), ph);
  }
  // the stage, transposed through LDS (thread (q0, cl) takes the points
  // ix = q0 + q T of the workgroup's line cl)
  const int cl = threadIdx.x % C::CPB, q0 = threadIdx.x / C::CPB;
  const long long sraw = (long long)blockIdx.x * C::CPB + cl;
  const bool sact = sraw < nlines;
  const long long sline = sact ? sraw : 0;
  const long long sinst = sline / inner, sli = sline - sinst * inner;
  const long long sb0 = sinst * a.nr + sli;
  const float* tb = stage_table(a, sinst / a.channels, sinst % a.channels);
#pragma unroll
  for (int s = 0; s < PT; ++s) lds[j + s * T] = v[0][s];
  __syncthreads();
  with_stage_mode(a.mode, [&](auto mc) {
    constexpr int M = decltype(mc)::value;
    constexpr bool first = !(M == ST2B || M == ST3B || M == ST3C || M == ST4B || M == ST4C || M == ST4D);
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const int ix = q0 + q * T;
      const cf r = lds_x[cl * LDSN + ix];
      const long long k = sli + (long long)ix * inner;
      const long long i = sb0 + (long long)ix * inner;
      if (!sact) continue;
      cf sv;
      if constexpr (first) {
        N0out[i] = r;
        sv = stage_core<M>(a.e_im, a.nr, tb, k, i, U0, S, S1, A, r, cf{0.f, 0.f});
      } else {
        sv = stage_core<M>(a.e_im, a.nr, tb, k, i, U0, S, S1, A, stage_reads_n0(M) ? N0[i] : cf{0.f, 0.f}, r);
      }
      if constexpr (EMIT) lds_x[cl * LDSN + ix] = sv;  // the slot this thread alone read
    }
  });
  if constexpr (EMIT) {
    // the new state's inverse x lines (times 1/N, every kz plane: the first
    // pass of irfftn) to K, from the lane groups' line order
    __syncthreads();
    cf e[1][PT];
#pragma unroll
    for (int s = 0; s < PT; ++s) e[0][s] = lds[j + s * T];
    cfft<N, PT, true, 1>(e, ln, lds, tr);  // only this group uses its LDS line until the next barrier
    // stored transposed through LDS, as the stage's accesses
    constexpr float rn = 1.f / (float)N;
#pragma unroll
    for (int s = 0; s < PT; ++s) lds[j + s * T] = rn * e[0][s];
    __syncthreads();
    if (sact) {
#pragma unroll
      for (int q = 0; q < PT; ++q) {
        const int ix = q0 + q * T;
        K[sb0 + (long long)ix * inner] = lds_x[cl * LDSN + ix];
      }
    }
  }
}

// Fused forward x lines, combine and stage of the kinds with fused z rows;
// hipErrorInvalidValue for the other kinds and sizes. K: where the new
// state's inverse x lines go (the emitting last stage of a rollout step), or
// nullptr.
hipError_t launch_xstage3d(const NL3DArgs& na, const Stage3DArgs& a, const cf* P, cf* N0out, int kcut, const cf* tw,
                           hipStream_t st, cf* K) {
  const bool first = a.mode == ST1 || a.mode == ST2A || a.mode == ST3A || a.mode == ST4A;
  if (first != (N0out != nullptr)) return hipErrorInvalidValue;  // the first stage keeps N0, the others read it
  auto go = [&](auto kern, int cpb) {
    const long long nlines = a.ninst * (long long)na.N * na.NZ;
    hipLaunchKernelGGL(kern, dim3((unsigned)((nlines + cpb - 1) / cpb)), dim3(256), 0, st, na, a, P, N0out, kcut,
                       tw, a.U0, a.N0, a.S, a.S1, a.A, K);
    return hipGetLastError();
  };
#define EXMI_XSK(NN, KK)                                                                              \
  case KK:                                                                                            \
    return K ? go(xstage3d_kernel<NN, KK, true>, XStageCfg<NN>::CPB)                                  \
             : go(xstage3d_kernel<NN, KK, false>, XStageCfg<NN>::CPB);
#define EXMI_XSTAGE(NN)                    \
  case NN:                                 \
    switch (na.kind) {                     \
      EXMI_XSK(NN, NL_GRADNORM)            \
      EXMI_XSK(NN, NL_GENERAL)             \
      EXMI_XSK(NN, NL_POLY)                \
      EXMI_XSK(NN, NL_GRAY_SCOTT)          \
      EXMI_XSK(NN, NL_BZ)                  \
      EXMI_XSK(NN, NL_CONV_CONS)           \
      EXMI_XSK(NN, NL_CONV_NONCONS)        \
      default: return hipErrorInvalidValue; \
    }
  switch (na.N) {
    EXMI_XSTAGE(16)
    EXMI_XSTAGE(32)
    EXMI_XSTAGE(64)
    EXMI_XSTAGE(128)
    default: return hipErrorInvalidValue;
  }
#undef EXMI_XSTAGE
#undef EXMI_XSK
}

hipError_t launch_stage_nl3d(const NL3DArgs& na, const Stage3DArgs& a, const cf* PH, cf* N0out, hipStream_t st) {
  if (na.kind == NL_PROJ3D && a.channels == 3) {
    hipLaunchKernelGGL(stage_nl3d_proj_kernel, dim3(grid_of(a.ninst / 3 * a.nr)), dim3(256), 0, st, na, a, PH,
                       N0out);
  } else {
    hipLaunchKernelGGL(stage_nl3d_kernel, dim3(grid_of(a.ninst * a.nr)), dim3(256), 0, st, na, a, PH, N0out);
  }
  return hipGetLastError();
}

}  // namespace exmi

namespace exmi {

// The wave equation's exact step in Fourier space (stepper/_wave.py:130-197)
// on (batch, 2, Nr) spectra of (h, v) in D = 1, 2, 3. The reference rotates
// into the modes (w +- v)/sqrt 2 with w = i c |k| h, advances them by
// exp(+- i c |k| dt), rotates back and divides by i c |k|; multiplied out
// this is, with theta = c |k| dt,
//   h' = cos(theta) h + sin(theta) / (c |k|) v,  v' = -c |k| sin(theta) h + cos(theta) v,
// and at the mean mode h' = h + dt v, v' = v (the reference's drift term).
// The product form avoids the float32 cancellation of w e^{i theta} -
// w e^{-i theta} when c |k| |h| >> |v| (the reference's own float32 error
// grows with c |k|); the coefficients are formed in float64.
__global__ void wave_spec_kernel(cf* X, long long batch, int D, int N, double kscale, double c, double dt) {
  const int NZ = N / 2 + 1;
  long long nr = NZ;
  for (int d = 1; d < D; ++d) nr *= N;
  const long long total = batch * nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long b = idx / nr, k = idx % nr;
    long long r = k / NZ;
    const int kz = (int)(k % NZ);
    double k2 = (double)kz * kz;
    for (int d = 1; d < D; ++d) {
      const int i = (int)(r % N);
      r /= N;
      const int kk = i < N / 2 ? i : i - N;
      k2 += (double)kk * kk;
    }
    const double ck = c * kscale * sqrt(k2);
    double sn, cs;
    sincos(ck * dt, &sn, &cs);
    const bool dc = k2 == 0.0;
    const float a = (float)cs, bh = (float)(dc ? dt : sn / ck), bv = (float)(-ck * sn);
    cf* h = X + (b * 2) * nr + k;
    cf* v = X + (b * 2 + 1) * nr + k;
    const cf h0 = *h, v0 = *v;
    *h = cf{fmaf(a, h0.re, bh * v0.re), fmaf(a, h0.im, bh * v0.im)};
    *v = cf{fmaf(bv, h0.re, a * v0.re), fmaf(bv, h0.im, a * v0.im)};
  }
}

hipError_t launch_wave_spec(cf* X, long long batch, int D, int N, double kscale, double c, double dt,
                            hipStream_t st) {
  long long nr = N / 2 + 1;
  for (int d = 1; d < D; ++d) nr *= N;
  hipLaunchKernelGGL(wave_spec_kernel, dim3(grid_of(batch * nr)), dim3(256), 0, st, X, batch, D, N, kscale, c, dt);
  return hipGetLastError();
}

}  // namespace exmi

namespace exmi {

// Poisson solve in Fourier space (_poisson.py): u_hat = -f_hat / Laplace_p(k),
// Laplace_p = sum_d (i kappa_d)^p (p even), and 0 where that operator is 0
// (the mean-zero solution). count spectra of N^(D-1) (N/2+1) modes, in place.
__global__ void poisson_spec_kernel(cf* X, long long count, int D, int N, double kscale, int order) {
  const int NZ = N / 2 + 1;
  long long nr = NZ;
  for (int d = 1; d < D; ++d) nr *= N;
  const long long total = count * nr;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long k = idx % nr;
    long long r = k / NZ;
    const int kz = (int)(k % NZ);
    // (i kappa)^p = (-1)^(p/2) kappa^p for even p
    const double sgn = (order / 2) % 2 ? -1.0 : 1.0;
    double op = sgn * pow(kscale * kz, order);
    for (int d = 1; d < D; ++d) {
      const int i = (int)(r % N);
      r /= N;
      const int kk = i < N / 2 ? i : i - N;
      op += sgn * pow(kscale * kk, order);
    }
    const float f = op == 0.0 ? 0.f : (float)(-1.0 / op);
    X[idx] = f * X[idx];
  }
}

hipError_t launch_poisson_spec(cf* X, long long count, int D, int N, double kscale, int order, hipStream_t st) {
  long long nr = N / 2 + 1;
  for (int d = 1; d < D; ++d) nr *= N;
  hipLaunchKernelGGL(poisson_spec_kernel, dim3(grid_of(count * nr)), dim3(256), 0, st, X, count, D, N, kscale, order);
  return hipGetLastError();
}

// out = x + a y over n floats (the forcing of ForcedStepper, _forced_stepper.py:33-56)
__global__ void axpy_kernel(long long n, float a, const float* __restrict__ x, const float* __restrict__ y,
                            float* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = fmaf(a, y[i], x[i]);
}

hipError_t launch_axpy(long long n, float a, const float* x, const float* y, float* out, hipStream_t st) {
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_of(n)), dim3(256), 0, st, n, a, x, y, out);
  return hipGetLastError();
}

}  // namespace exmi
