// Register/LDS Stockham FFT building blocks for gfx950 (wave64).
//
// A complex transform of M points is owned by T lanes; lane j holds the PT =
// M/T points j + s*T (s = 0..PT-1) in registers ("lane-strided order"). The
// Stockham passes (Govindaraju et al., SC'08 formulation) read that order and
// the last pass produces natural-order output again in the same lane-strided
// registers, so transforms chain without any reordering: a pass of radix R
// and stride NS reads x[b + r*M/R] and writes y[(b/NS)*NS*R + b%NS + r*NS]
// after the twiddle W_{NS*R}^{r*(b%NS)}. Between passes the data crosses LDS
// once (one store, one barrier, one load).
//
// Real transforms of N = 2M points use the half-length packing
// z[n] = x[2n] + i x[2n+1] with the usual even/odd split post-processing;
// the partner mode M-k lives in another lane and is fetched through LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace exmi {

struct cf {
  float re, im;
};

__device__ __forceinline__ cf mk(float r, float i) { return cf{r, i}; }
__device__ __forceinline__ cf operator+(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cf operator-(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cf operator*(float s, cf a) { return cf{s * a.re, s * a.im}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
  return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cf cmulc(cf a, cf b) {  // a * conj(b)
  return cf{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
__device__ __forceinline__ cf conj(cf a) { return cf{a.re, -a.im}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ cf rot90(cf a) {
  return INV ? cf{-a.im, a.re} : cf{a.im, -a.re};
}

// ---------------------------------------------------------------- butterflies
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(cf* v) {
    cf a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <bool INV>
struct Dft<4, INV> {
  __device__ __forceinline__ static void run(cf* v) {
    cf t0 = v[0] + v[2], t1 = v[0] - v[2];
    cf t2 = v[1] + v[3], t3 = rot90<INV>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
  }
};

template <bool INV>
struct Dft<8, INV> {
  __device__ __forceinline__ static void run(cf* v) {
    const float h = 0.70710678118654752440f;
    cf e[4] = {v[0], v[2], v[4], v[6]};
    cf o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    // o[k] *= W8^k (forward W8 = exp(-i pi/4); inverse conjugate)
    cf o1 = o[1], o3 = o[3];
    if (!INV) {
      o[1] = cf{h * (o1.re + o1.im), h * (o1.im - o1.re)};
      o[3] = cf{h * (o3.im - o3.re), -h * (o3.re + o3.im)};
    } else {
      o[1] = cf{h * (o1.re - o1.im), h * (o1.im + o1.re)};
      o[3] = cf{-h * (o3.re + o3.im), h * (o3.re - o3.im)};
    }
    o[2] = rot90<INV>(o[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = e[k] + o[k];
      v[k + 4] = e[k] - o[k];
    }
  }
};

template <bool INV>
struct Dft<16, INV> {
  // 4 x 4 decomposition: column DFT4s, inner twiddles W16^(r*c), row DFT4s.
  __device__ __forceinline__ static void run(cf* v) {
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    cf a[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      cf t[4] = {v[c], v[c + 4], v[c + 8], v[c + 12]};
      Dft<4, INV>::run(t);
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r][c] = t[r];
    }
    // twiddle W16^(r*c), forward exp(-2 pi i r c / 16)
    const float sg = INV ? 1.f : -1.f;
    const cf w1 = cf{c1, sg * s1}, w2 = cf{h, sg * h}, w3 = cf{s1, sg * c1};
    const cf w6 = cf{-h, sg * h}, w9 = cf{-c1, -sg * s1};
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = rot90<INV>(a[2][2]);
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      cf t[4] = {a[r][0], a[r][1], a[r][2], a[r][3]};
      Dft<4, INV>::run(t);
#pragma unroll
      for (int c = 0; c < 4; ++c) v[r + 4 * c] = t[c];
    }
  }
};

// --------------------------------------------------------------- pass plans
// Radix sequence for M points with PT points per lane: radix-PT passes while
// they fit, then one smaller pass for the remainder.
__host__ __device__ constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }

template <int M, int PT>
struct Plan {
  static constexpr int LM = ilog2(M);
  static constexpr int LR = ilog2(PT);
  static constexpr int NFULL = LM / LR;
  static constexpr int REM = LM - NFULL * LR;  // log2 of the last radix (0 = none)
  static constexpr int NPASS = NFULL + (REM ? 1 : 0);
  __host__ __device__ static constexpr int radix(int p) { return p < NFULL ? PT : (1 << REM); }
  __host__ __device__ static constexpr int stride(int p) {  // NS of pass p
    return p == 0 ? 1 : stride(p - 1) * radix(p - 1);
  }
};

// LDS index padding: one pad slot every 16 complex entries.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
template <int M>
struct LdsSize {
  static constexpr int value = M + (M >> 4) + 1;
};

// One Stockham pass on the lane's registers. `v` holds the lane's PT inputs
// in lane-strided order (input x[j + s*T] in v[s]) and receives the outputs
// y[j + s*T] in v[s] after the exchange through `lds`, except for the last
// pass which needs no exchange.
// `tw` is the table exp(-2 pi i m / M), m = 0..M-1.
template <int M, int T, int R, int NS, bool INV, bool LAST>
__device__ __forceinline__ void stockham_pass(cf* v, int j, cf* lds, const cf* __restrict__ tw) {
  constexpr int PT = M / T;
  constexpr int NB = PT / R;   // butterflies per lane
  constexpr int MR = M / R;    // butterfly count
  constexpr int TWS = M / (NS * R);
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = j + q * T;
    cf x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[q + r * NB];  // x[b + r*MR] = v[(b + r*MR - j)/T]
    if constexpr (NS > 1) {
      const int k = b % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cf w = tw[(r * k * TWS) & (M - 1)];
        x[r] = INV ? cmulc(x[r], w) : cmul(x[r], w);
      }
    }
    Dft<R, INV>::run(x);
    if constexpr (LAST) {
      // NS*R == M: output index b + r*NS = j + (q + r*NB)*T
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * NB] = x[r];
    } else {
      const int base = (b / NS) * NS * R + (b % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) lds[pad(base + r * NS)] = x[r];
    }
  }
  (void)MR;
}

template <int M, int T>
__device__ __forceinline__ void lds_load(cf* v, int j, const cf* lds) {
  constexpr int PT = M / T;
#pragma unroll
  for (int s = 0; s < PT; ++s) v[s] = lds[pad(j + s * T)];
}

// Sync among the lanes sharing one transform's LDS region. When a transform
// fits inside one wavefront the exchange is wave-local.
template <int T>
__device__ __forceinline__ void xsync() {
  if constexpr (T <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <int M, int PT, bool INV, int P>
struct PassRunner {
  using PL = Plan<M, PT>;
  static constexpr int T = M / PT;
  __device__ __forceinline__ static void run(cf* v, int j, cf* lds, const cf* __restrict__ tw) {
    constexpr bool LAST = (P == PL::NPASS - 1);
    stockham_pass<M, T, PL::radix(P), PL::stride(P), INV, LAST>(v, j, lds, tw);
    if constexpr (!LAST) {
      xsync<T>();
      lds_load<M, T>(v, j, lds);
      xsync<T>();
      PassRunner<M, PT, INV, P + 1>::run(v, j, lds, tw);
    }
  }
};

// Complex FFT of M points, in place on lane-strided registers.
template <int M, int PT, bool INV>
__device__ __forceinline__ void cfft(cf* v, int j, cf* lds, const cf* __restrict__ tw) {
  PassRunner<M, PT, INV, 0>::run(v, j, lds, tw);
}

// ----------------------------------------------------------- real transforms
// Spectrum layout: lane j, slot s holds X[j + s*T] (k = 0..M-1); `nyq` holds
// X[M] (meaningful in lane 0 only, kept in every lane of the group).
//
// r2c: z (packed even/odd samples, lane-strided) -> X. `twr` is the table
// exp(-2 pi i k / N), k = 0..M.
template <int M, int PT>
__device__ __forceinline__ void rfft(cf* z, cf& nyq, int j, cf* lds, const cf* __restrict__ tw,
                                     const cf* __restrict__ twr) {
  constexpr int T = M / PT;
  cfft<M, PT, false>(z, j, lds, tw);
  // partner exchange Z[M-k]
#pragma unroll
  for (int s = 0; s < PT; ++s) lds[pad(j + s * T)] = z[s];
  xsync<T>();
  cf zm[PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int k = j + s * T;
    zm[s] = lds[pad((M - k) & (M - 1))];
  }
  xsync<T>();
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int k = j + s * T;
    const cf a = z[s], b = conj(zm[s]);
    const cf e = 0.5f * (a + b);               // even-sample spectrum
    const cf d = a - b;                        // 2i * odd-sample spectrum
    const cf o = cf{0.5f * d.im, -0.5f * d.re};  // d / (2i)
    const cf w = twr[k];
    z[s] = e + cmul(w, o);
  }
  // k = 0 (lane 0, slot 0): zm[0] = Z[0]; X[0] = Re Z0 + Im Z0 and the
  // Nyquist mode X[M] = Re Z0 - Im Z0, both real. `nyq` is meaningful in
  // lane 0 only.
  nyq = cf{zm[0].re - zm[0].im, 0.f};
  if (j == 0) z[0] = cf{zm[0].re + zm[0].im, 0.f};
}

// c2r: X (lane-strided spectrum + nyq in lane 0) -> z packed physical samples
// multiplied by `scale` (1/N for the numpy irfft convention). Imaginary parts
// of X[0] and X[M] are ignored.
template <int M, int PT>
__device__ __forceinline__ void irfft(cf* z, cf nyq, int j, cf* lds, const cf* __restrict__ tw,
                                      const cf* __restrict__ twr, float scale) {
  constexpr int T = M / PT;
  if (j == 0) {
    // pack X[0] (real) and X[M] (real) so that the generic formula at k = 0
    // reads X[M-0] = X[M] from slot 0 after the exchange below.
    z[0] = cf{z[0].re, nyq.re};
  }
#pragma unroll
  for (int s = 0; s < PT; ++s) lds[pad(j + s * T)] = z[s];
  xsync<T>();
  cf xm[PT];
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int k = j + s * T;
    xm[s] = lds[pad((M - k) & (M - 1))];
  }
  xsync<T>();
#pragma unroll
  for (int s = 0; s < PT; ++s) {
    const int k = j + s * T;
    cf a = z[s], b = xm[s];
    if (k == 0) {
      // a = (X0, XM) packed; X[M-0] := XM
      const float x0 = a.re, xmv = a.im;
      a = cf{x0, 0.f};
      b = cf{xmv, 0.f};
    }
    b = conj(b);
    const cf e = a + b;                       // 2 * even spectrum
    const cf d = a - b;                       // 2 * W^-k * odd spectrum / W^-k
    const cf w = twr[k];                      // exp(-2 pi i k/N); need exp(+...)
    const cf o = cmulc(d, w);                 // d * exp(+2 pi i k / N)
    // Z = E + i O
    z[s] = scale * cf{e.re - o.im, e.im + o.re};
  }
  cfft<M, PT, true>(z, j, lds, tw);
}

}  // namespace exmi
