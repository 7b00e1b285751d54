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
// Twiddled passes multiply inputs 1..R-1 of every butterfly by twiddles w
// (conj(w) for the inverse). Where a first-stage radix-2 butterfly meets a
// twiddled input, a + b w is formed with four fused multiply-adds and
// a - b w as 2 a - (a + b w): 6 operations instead of the 8 of a complex
// multiply followed by an add and a subtract.
template <bool INV>
__device__ __forceinline__ cf twid(cf b, cf w) {
  return INV ? cmulc(b, w) : cmul(b, w);
}
// a + b w (a + b conj(w) for INV) in four fused multiply-adds
template <bool INV>
__device__ __forceinline__ cf cfma(cf a, cf b, cf w) {
  if constexpr (INV)  // b conj(w) = (b.re w.re + b.im w.im, b.im w.re - b.re w.im)
    return cf{fmaf(b.im, w.im, fmaf(b.re, w.re, a.re)), fmaf(-b.re, w.im, fmaf(b.im, w.re, a.im))};
  else  // b w = (b.re w.re - b.im w.im, b.re w.im + b.im w.re)
    return cf{fmaf(-b.im, w.im, fmaf(b.re, w.re, a.re)), fmaf(b.im, w.re, fmaf(b.re, w.im, a.im))};
}
template <bool INV>
__device__ __forceinline__ void bfly_tw(cf a, cf b, cf w, cf& s, cf& d) {
  s = cfma<INV>(a, b, w);
  d = cf{fmaf(2.f, a.re, -s.re), fmaf(2.f, a.im, -s.im)};
}

// DFT4 of (y0, y1, y2, y3) with y_r = x_r w_r; w0 = 1 unless TW0.
// The pairs (y0, y2) and (y1, y3) are the first-stage butterflies.
template <bool INV, bool TW0>
__device__ __forceinline__ void dft4_tw(cf* v, cf w0, cf w1, cf w2, cf w3) {
  cf t0, t1, t2, t3;
  bfly_tw<INV>(TW0 ? twid<INV>(v[0], w0) : v[0], v[2], w2, t0, t1);
  bfly_tw<INV>(twid<INV>(v[1], w1), v[3], w3, t2, t3);
  t3 = rot90<INV>(t3);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(cf* v) {
    cf a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
  // w[r - 1] multiplies input r
  __device__ __forceinline__ static void run_tw(cf* v, const cf* w) { bfly_tw<INV>(v[0], v[1], w[0], v[0], v[1]); }
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
  __device__ __forceinline__ static void run_tw(cf* v, const cf* w) {
    dft4_tw<INV, false>(v, cf{1.f, 0.f}, w[0], w[1], w[2]);
  }
};

template <bool INV>
struct Dft<3, INV> {
  // X1, X2 = a - (b + c)/2 -+ i s (b - c) (forward), s = sin(2 pi / 3)
  __device__ __forceinline__ static void run(cf* v) {
    const float s3 = 0.86602540378443864676f;
    const cf a = v[0], t1 = v[1] + v[2], d = v[1] - v[2];
    const cf t2 = cf{fmaf(-0.5f, t1.re, a.re), fmaf(-0.5f, t1.im, a.im)};
    const cf rd = rot90<INV>(cf{s3 * d.re, s3 * d.im});  // -i s d forward, +i s d inverse
    v[0] = a + t1;
    v[1] = t2 + rd;
    v[2] = t2 - rd;
  }
  __device__ __forceinline__ static void run_tw(cf* v, const cf* w) {
    v[1] = twid<INV>(v[1], w[0]);
    v[2] = twid<INV>(v[2], w[1]);
    run(v);
  }
};

template <bool INV>
struct Dft<5, INV> {
  // t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3;
  // X1,4 = x0 + c1 t1 + c2 t2 -+ i (s1 t3 + s2 t4),
  // X2,3 = x0 + c2 t1 + c1 t2 -+ i (s2 t3 - s1 t4) (forward; + - for inverse),
  // c_j = cos(2 pi j / 5), s_j = sin(2 pi j / 5).
  __device__ __forceinline__ static void run(cf* v) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const cf x0 = v[0], t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
    const cf a1 = cf{fmaf(c1, t1.re, fmaf(c2, t2.re, x0.re)), fmaf(c1, t1.im, fmaf(c2, t2.im, x0.im))};
    const cf a2 = cf{fmaf(c2, t1.re, fmaf(c1, t2.re, x0.re)), fmaf(c2, t1.im, fmaf(c1, t2.im, x0.im))};
    const cf b1 = rot90<INV>(cf{fmaf(s1, t3.re, s2 * t4.re), fmaf(s1, t3.im, s2 * t4.im)});
    const cf b2 = rot90<INV>(cf{fmaf(s2, t3.re, -s1 * t4.re), fmaf(s2, t3.im, -s1 * t4.im)});
    v[0] = x0 + t1 + t2;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
  }
  __device__ __forceinline__ static void run_tw(cf* v, const cf* w) {
#pragma unroll
    for (int r = 1; r < 5; ++r) v[r] = twid<INV>(v[r], w[r - 1]);
    run(v);
  }
};

template <bool INV>
struct Dft<8, INV> {
  // radix-2 split into two DFT4s; the odd outputs' W8^1 and W8^3 factors
  // (h (1 -+ i), h (-1 -+ i)) are folded into fused multiply-adds of the
  // final butterfly.
  __device__ __forceinline__ static void run(cf* v) {
    cf e[4] = {v[0], v[2], v[4], v[6]};
    cf o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    combine(e, o, v);
  }
  __device__ __forceinline__ static void run_tw(cf* v, const cf* w) {
    cf e[4] = {v[0], v[2], v[4], v[6]};
    cf o[4] = {v[1], v[3], v[5], v[7]};
    dft4_tw<INV, false>(e, cf{1.f, 0.f}, w[1], w[3], w[5]);
    dft4_tw<INV, true>(o, w[0], w[2], w[4], w[6]);
    combine(e, o, v);
  }
  __device__ __forceinline__ static void combine(const cf* e, const cf* o, cf* v) {
    const float h = 0.70710678118654752440f;
    // forward: W8 o1 = h ((o1.re + o1.im) + i (o1.im - o1.re)),
    //          W8^3 o3 = h ((o3.im - o3.re) - i (o3.re + o3.im));
    // inverse: conjugate factors.
    const float p1 = INV ? (o[1].re - o[1].im) : (o[1].re + o[1].im);
    const float q1 = INV ? (o[1].im + o[1].re) : (o[1].im - o[1].re);
    const float p3 = INV ? -(o[3].re + o[3].im) : (o[3].im - o[3].re);
    const float q3 = INV ? (o[3].re - o[3].im) : -(o[3].re + o[3].im);
    const cf o2 = rot90<INV>(o[2]);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = cf{fmaf(h, p1, e[1].re), fmaf(h, q1, e[1].im)};
    v[5] = cf{fmaf(-h, p1, e[1].re), fmaf(-h, q1, e[1].im)};
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = cf{fmaf(h, p3, e[3].re), fmaf(h, q3, e[3].im)};
    v[7] = cf{fmaf(-h, p3, e[3].re), fmaf(-h, q3, e[3].im)};
  }
};

template <bool INV>
struct Dft<16, INV> {
  // 4 x 4 decomposition: column DFT4s, inner twiddles W16^(r*c), row DFT4s.
  __device__ __forceinline__ static void run(cf* v) { run_impl<false>(v, nullptr); }
  __device__ __forceinline__ static void run_tw(cf* v, const cf* w) { run_impl<true>(v, w); }
  template <bool TW>
  __device__ __forceinline__ static void run_impl(cf* v, const cf* w) {
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    cf a[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      cf t[4] = {v[c], v[c + 4], v[c + 8], v[c + 12]};
      if constexpr (TW) {
        const cf one{1.f, 0.f};
        if (c == 0) dft4_tw<INV, false>(t, one, w[3], w[7], w[11]);
        else dft4_tw<INV, true>(t, w[c - 1], w[c + 3], w[c + 7], w[c + 11]);
      } else {
        Dft<4, INV>::run(t);
      }
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

// Single-lane transforms of the non-power-of-two sizes N = 48 and N = 50
// (every point of a transform in one lane's registers, T = 1): 24 = 8 x 3,
// 25 = 5 x 5.
template <>
struct Plan<24, 24> {
  static constexpr int NPASS = 2;
  __host__ __device__ static constexpr int radix(int p) { return p == 0 ? 8 : 3; }
  __host__ __device__ static constexpr int stride(int p) { return p == 0 ? 1 : 8; }
};
template <>
struct Plan<25, 25> {
  static constexpr int NPASS = 2;
  __host__ __device__ static constexpr int radix(int) { return 5; }
  __host__ __device__ static constexpr int stride(int p) { return p == 0 ? 1 : 5; }
};

// LDS index padding: one pad slot every 16 complex entries.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
template <int M>
struct LdsSize {
  static constexpr int value = M + (M >> 4) + 1;
};

// Per-lane LDS addressing. With pad(x) = x + x/16 and power-of-two sizes,
// pad(base + c) = pad(base) + c + c/16 for every offset c the passes use
// (c a multiple of the pass stride below one 16-slot line, or of 16), so each
// access is one per-lane base register plus an immediate offset.
__host__ __device__ constexpr int cpad(int c) { return c + c / 16; }

template <int M, int PT>
struct Lane {
  static constexpr int T = M / PT;
  int j;    // lane index within the transform
  int rd;   // pad(j): base of the lane-strided reads j + s*T
  int mir;  // pad(T - j): base of the mirrored reads M - j - s*T
  __device__ __forceinline__ explicit Lane(int jj) : j(jj), rd(pad(jj)), mir(pad(T - jj)) {}
  // LDS index of element j + s*T
  __device__ __forceinline__ int at(int s) const { return rd + cpad(s * T); }
  // LDS index of element (M - (j + s*T)) mod M; lane 0 mirrors onto its
  // own slots (PT - s) mod PT.
  __device__ __forceinline__ int mirror(int s) const {
    return j == 0 ? cpad(((PT - s) % PT) * T) : mir + cpad((PT - 1 - s) * T);
  }
};

// One Stockham pass on the lanes' registers, for NT independent transforms
// handled together (their instruction streams interleave). v[n] holds the
// lane's PT inputs of transform n in lane-strided order (x[j + s*T] in
// v[n][s]) and receives the outputs y[j + s*T]. All NT transforms exchange
// through the same LDS buffer, one after the other (write n, read n, write
// n+1, ...): LDS operations of a wavefront complete in program order, so
// transform n+1's stores cannot overtake transform n's loads. The last pass
// needs no exchange. `pw` are the lane's twiddles of this pass (TwRegs).
template <int T>
__device__ __forceinline__ void xsync();

template <int M, int PT, int R, int NS, bool INV, bool LAST, int NT>
__device__ __forceinline__ void stockham_pass(cf (*v)[PT], const Lane<M, PT>& ln, cf* lds, const cf* pw) {
  constexpr int T = M / PT;
  constexpr int NB = PT / R;  // butterflies per lane
  static_assert(NT == 1 || T <= 64, "interleaved transforms must live in one wavefront");
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      cf x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[n][q + r * NB];  // x[b + r*M/R], b = j + q*T
      if constexpr (NS > 1) Dft<R, INV>::run_tw(x, pw + q * (R - 1));
      else Dft<R, INV>::run(x);
#pragma unroll
      for (int r = 0; r < R; ++r) v[n][q + r * NB] = x[r];
    }
  }
  if constexpr (!LAST && T == 1) {
    // one lane holds the whole transform: the exchange is a compile-time
    // permutation of its registers
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      cf y[PT];
#pragma unroll
      for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) y[(q / NS) * NS * R + (q % NS) + r * NS] = v[n][q + r * NB];
#pragma unroll
      for (int s = 0; s < PT; ++s) v[n][s] = y[s];
    }
  } else if constexpr (!LAST) {
    // output y[(b/NS)*NS*R + b%NS + r*NS] = v[n][q + r*NB]
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int b = ln.j + q * T;
        const int wb = pad((b / NS) * NS * R + (b % NS));
#pragma unroll
        for (int r = 0; r < R; ++r) lds[wb + cpad(r * NS)] = v[n][q + r * NB];
      }
      xsync<T>();
#pragma unroll
      for (int s = 0; s < PT; ++s) v[n][s] = lds[ln.at(s)];
      xsync<T>();
    }
  }
  // LAST: NS*R == M, output index b + r*NS = j + (q + r*NB)*T is already the
  // lane-strided slot q + r*NB.
}

// Sync among the lanes sharing one transform's LDS region. When a transform
// fits inside one wavefront the exchange is wave-local (LDS operations of a
// wave complete in order; the fences only stop compiler reordering).
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

template <int M, int PT>
struct TwRegs;

template <int M, int PT, bool INV, int P, int NT>
struct PassRunner {
  using PL = Plan<M, PT>;
  static constexpr int T = M / PT;
  __device__ __forceinline__ static void run(cf (*v)[PT], const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& tr) {
    constexpr bool LAST = (P == PL::NPASS - 1);
    stockham_pass<M, PT, PL::radix(P), PL::stride(P), INV, LAST, NT>(v, ln, lds, tr.pw[P]);
    if constexpr (!LAST) PassRunner<M, PT, INV, P + 1, NT>::run(v, ln, lds, tr);
  }
};

// Complex FFTs of M points, in place on lane-strided registers.
template <int M, int PT, bool INV, int NT>
__device__ __forceinline__ void cfft(cf (*v)[PT], const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& tr) {
  PassRunner<M, PT, INV, 0, NT>::run(v, ln, lds, tr);
}

// Table of the real-transform pre/post-processing, per mode k = 0..M-1
// (N = 2M, w_k = exp(-2 pi i k / N)): fw[k] = -i w_k. The transforms carry
// power-of-two scale factors (exact in floating point) so that one table
// serves both directions:
//   forward  X2[k] = (Z[k] + Z*[M-k]) + fw[k] (Z[k] - Z*[M-k])  = 2 X[k]
//   inverse  Z'[k] = (X[k] + X*[M-k]) + conj(fw[k]) (X[k] - X*[M-k])
//            and the M-point inverse DFT of Z' gives N * irfft(X).
struct RealTables {
  const cf* tw;  // exp(-2 pi i m / M), m < M
  const cf* fw;  // -i exp(-2 pi i k / N), k < M
};

// Per-lane twiddle registers: every twiddle a lane multiplies by in any pass,
// and its pre/post-processing factors, loaded once per kernel so that the
// transforms themselves issue no memory loads.
template <int M, int PT>
struct TwRegs {
  using PL = Plan<M, PT>;
  static constexpr int T = M / PT;
  cf pw[PL::NPASS][PT];  // pass p, entry q*(R-1) + (r-1); pass 0 needs none
  cf fw[PT];
  __device__ __forceinline__ void load(int j, const RealTables& rt) {
    load_pass<1>(j, rt.tw);
#pragma unroll
    for (int s = 0; s < PT; ++s) fw[s] = rt.fw[j + s * T];
  }
  template <int P>
  __device__ __forceinline__ void load_pass(int j, const cf* __restrict__ tw) {
    if constexpr (P < PL::NPASS) {
      constexpr int R = PL::radix(P), NS = PL::stride(P), NB = PT / R, TWS = M / (NS * R);
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int k = (j + q * T) % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) pw[P][q * (R - 1) + r - 1] = tw[r * k * TWS];
      }
      load_pass<P + 1>(j, tw);
    }
  }
};

// ----------------------------------------------------------- real transforms
// Spectrum layout: lane j, slot s holds X[j + s*T] (k = 0..M-1); nyq holds
// X[M] (meaningful in lane 0 only).
//
// r2c: z (packed even/odd samples x[2n] + i x[2n+1], lane-strided) -> 2 X.
// Modes above `kcut` are not formed (set to zero): callers that dealias the
// result pass the dealias cutoff; slots entirely above it are skipped with a
// uniform branch.
template <int M, int PT, int NT>
__device__ __forceinline__ void rfft(cf (*z)[PT], cf* nyq, const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& tr,
                                     int kcut = M) {
  constexpr int T = M / PT;
  cfft<M, PT, false, NT>(z, ln, lds, tr);
  cf zm[NT][PT];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if constexpr (T == 1) {
#pragma unroll
      for (int s = 0; s < PT; ++s) zm[n][s] = z[n][(PT - s) % PT];
    } else {
#pragma unroll
      for (int s = 0; s < PT; ++s) lds[ln.at(s)] = z[n][s];
      xsync<T>();
#pragma unroll
      for (int s = 0; s < PT; ++s) zm[n][s] = lds[ln.mirror(s)];
      xsync<T>();
    }
  }
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    // lane 0 slot 0 (k = 0): zm[0] = Z[0]; 2 X[M] = 2 (Re Z0 - Im Z0)
    nyq[n] = cf{2.f * (zm[n][0].re - zm[n][0].im), 0.f};
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      if (s * T > kcut) {
        z[n][s] = cf{0.f, 0.f};
        continue;
      }
      const cf a = z[n][s], b = conj(zm[n][s]);
      // at k = 0 this yields 2 X[0] = 2 (Re Z0 + Im Z0) since fw[0] = -i
      z[n][s] = cfma<false>(a + b, a - b, tr.fw[s]);
    }
  }
}

// c2r: X (lane-strided spectrum + nyq in lane 0) -> z packed physical samples
// times N (numpy's irfft without its 1/N). Imaginary parts of X[0] and X[M]
// are ignored.
template <int M, int PT, int NT>
__device__ __forceinline__ void irfft(cf (*z)[PT], const cf* nyq, const Lane<M, PT>& ln, cf* lds,
                                      const TwRegs<M, PT>& tr) {
  constexpr int T = M / PT;
  // pack X[0] (real) and X[M] (real) so that the generic formula at k = 0
  // reads X[M] as the partner of X[0] after the exchange below
  if (ln.j == 0) {
#pragma unroll
    for (int n = 0; n < NT; ++n) z[n][0] = cf{z[n][0].re, nyq[n].re};
  }
  cf xm[NT][PT];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if constexpr (T == 1) {
#pragma unroll
      for (int s = 0; s < PT; ++s) xm[n][s] = z[n][(PT - s) % PT];
    } else {
#pragma unroll
      for (int s = 0; s < PT; ++s) lds[ln.at(s)] = z[n][s];
      xsync<T>();
#pragma unroll
      for (int s = 0; s < PT; ++s) xm[n][s] = lds[ln.mirror(s)];
      xsync<T>();
    }
  }
  if (ln.j == 0) {
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      xm[n][0] = cf{z[n][0].im, 0.f};
      z[n][0] = cf{z[n][0].re, 0.f};
    }
  }
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int s = 0; s < PT; ++s) {
      const cf a = z[n][s], b = conj(xm[n][s]);
      z[n][s] = cfma<true>(a + b, a - b, tr.fw[s]);
    }
  cfft<M, PT, true, NT>(z, ln, lds, tr);
}

// ------------------------------------------------------ spectral layouts
// Where the M+1 modes of a real transform's spectrum live (T lanes x PT
// slots + one extra register per lane, meaningful in lane 0):
//   NATURAL: slot s of lane j holds mode j + s*T; the extra holds X[M].
//   PAIRED:  slots s < H = PT/2 hold k = j + s*T (k < M/2) and slot H + h
//            holds the partner M - (j + h*T); the extra holds X[M/2]. Lane 0
//            slot H is the Nyquist mode M.
// The paired layout keeps every mode next to its partner M - k, which the
// real pre/post-processing combines: one formula then yields both modes of a
// pair, and the exchange with the partner lane moves half the data.
enum SpecLayout { NATURAL = 0, PAIRED = 1 };

template <int M, int PT, int LAY>
struct Modes {
  static constexpr int T = M / PT;
  static constexpr int H = PT / 2;
  static constexpr bool lower(int s) { return LAY == NATURAL || s < H; }
  // mode of slot s in lane j
  __device__ __forceinline__ static int k(int j, int s) { return lower(s) ? j + s * T : M - j - (s - H) * T; }
  // smallest / largest mode of slot s over the lanes (uniform across the wave)
  __host__ __device__ static constexpr int klo(int s) { return lower(s) ? s * T : M - (s - H) * T - (T - 1); }
  __host__ __device__ static constexpr int khi(int s) { return lower(s) ? s * T + T - 1 : M - (s - H) * T; }
  static constexpr int kextra = LAY == NATURAL ? M : M / 2;
};

// r2c into the paired layout: z (packed samples, lane-strided) -> 2 X, and
// xm = 2 X[M/2] (lane 0; an unused value elsewhere). Modes above `kcut` are not formed
// where a whole slot lies above it.
//   2 X[k]   = s + P,   2 X[M-k] = conj(s - P) = conj(2 s - (s + P)),
//   s = Z[k] + Z*[M-k], P = fw[k] (Z[k] - Z*[M-k]),
// the second from fw[M-k] = conj(fw[k]); at k = 0 the pair is (X[0], X[M])
// with Z[0] its own partner, and 2 X[M/2] = 2 conj(Z[M/2]) since
// fw[M/2] = -1.
template <int M, int PT, int NT>
__device__ __forceinline__ void rfft_paired(cf (*z)[PT], cf* xm, const Lane<M, PT>& ln, cf* lds,
                                            const TwRegs<M, PT>& tr, int kcut = M) {
  constexpr int T = M / PT, H = PT / 2;
  using MD = Modes<M, PT, PAIRED>;
  cfft<M, PT, false, NT>(z, ln, lds, tr);
  cf zp[NT][H];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if constexpr (T == 1) {
#pragma unroll
      for (int h = 0; h < H; ++h) zp[n][h] = z[n][(PT - h) % PT];
    } else {
#pragma unroll
      for (int s = H; s < PT; ++s) lds[ln.at(s)] = z[n][s];
      xsync<T>();
#pragma unroll
      for (int h = 0; h < H; ++h) zp[n][h] = lds[ln.mirror(h)];
      xsync<T>();
    }
  }
  // Lane-0 special cases are selects, not branches: a divergent branch would
  // split the interleaved instruction streams of the NT transforms.
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if (ln.j == 0) zp[n][0] = z[n][0];  // element 0 is its own partner and is not exchanged
    // meaningful in lane 0 (Z[M/2]); other lanes carry a finite unused value
    xm[n] = M / 2 <= kcut ? cf{2.f * z[n][H].re, -2.f * z[n][H].im} : cf{0.f, 0.f};
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const bool lo_drop = MD::klo(h) > kcut, hi_drop = MD::klo(H + h) > kcut;
      if (lo_drop && hi_drop) {
        z[n][h] = z[n][H + h] = cf{0.f, 0.f};
        continue;
      }
      const cf a = z[n][h], b = conj(zp[n][h]);
      if (hi_drop) {
        z[n][h] = cfma<false>(a + b, a - b, tr.fw[h]);
        z[n][H + h] = cf{0.f, 0.f};
      } else {
        cf x0, x1;
        bfly_tw<false>(a + b, a - b, tr.fw[h], x0, x1);
        z[n][h] = lo_drop ? cf{0.f, 0.f} : x0;
        z[n][H + h] = conj(x1);
      }
    }
  }
}

// c2r from the paired layout: X (+ xm = X[M/2] in lane 0) -> packed samples
// times N. Imaginary parts of X[0] and X[M] are ignored. `kin`: modes above
// it are zero on entry (a dealiased spectrum); pairs whose partner slot lies
// wholly above it use the shorter formula.
//   Z'[k] = s + P,  Z'[M-k] = conj(s - P),
//   s = X[k] + X*[M-k], P = conj(fw[k]) (X[k] - X*[M-k]).
template <int M, int PT, int NT>
__device__ __forceinline__ void irfft_paired(cf (*z)[PT], const cf* xm, const Lane<M, PT>& ln, cf* lds,
                                             const TwRegs<M, PT>& tr, int kin = M) {
  constexpr int T = M / PT, H = PT / 2;
  using MD = Modes<M, PT, PAIRED>;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if (ln.j == 0) {
      z[n][0].im = 0.f;
      z[n][H].im = 0.f;
    }
    cf send[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const cf a = z[n][h];
      cf x0, x1;
      if (MD::klo(H + h) > kin) {  // X[M-k] = 0 for every lane
        bfly_tw<true>(a, a, tr.fw[h], x0, x1);
      } else {
        const cf b = conj(z[n][H + h]);
        bfly_tw<true>(a + b, a - b, tr.fw[h], x0, x1);
      }
      z[n][h] = x0;
      send[h] = conj(x1);
    }
    // Z'[M-k] to the natural slot of element M-k (lane 0, h = 0 writes the
    // meaningless Z'[M] to element 0's slot, which is not read back)
    if constexpr (T == 1) {
#pragma unroll
      for (int h = 1; h < H; ++h) z[n][PT - h] = send[h];
    } else {
#pragma unroll
      for (int h = 0; h < H; ++h) lds[ln.mirror(h)] = send[h];
      xsync<T>();
#pragma unroll
      for (int s = H; s < PT; ++s) z[n][s] = lds[ln.at(s)];
      xsync<T>();
    }
    if (ln.j == 0) z[n][H] = cf{2.f * xm[n].re, -2.f * xm[n].im};  // Z'[M/2], not exchanged
  }
  cfft<M, PT, true, NT>(z, ln, lds, tr);
}

// Layout dispatch.
template <int LAY, int M, int PT, int NT>
__device__ __forceinline__ void rfft_l(cf (*z)[PT], cf* x, const Lane<M, PT>& ln, cf* lds, const TwRegs<M, PT>& tr,
                                       int kcut = M) {
  if constexpr (LAY == PAIRED) rfft_paired<M, PT, NT>(z, x, ln, lds, tr, kcut);
  else rfft<M, PT, NT>(z, x, ln, lds, tr, kcut);
}
template <int LAY, int M, int PT, int NT>
__device__ __forceinline__ void irfft_l(cf (*z)[PT], const cf* x, const Lane<M, PT>& ln, cf* lds,
                                        const TwRegs<M, PT>& tr, int kin = M) {
  if constexpr (LAY == PAIRED) irfft_paired<M, PT, NT>(z, x, ln, lds, tr, kin);
  else irfft<M, PT, NT>(z, x, ln, lds, tr);
}

}  // namespace exmi
