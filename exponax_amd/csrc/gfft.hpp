// Runtime-size mixed-radix FFTs (gfft.hip) for grid sizes outside the fused
// kernels' compile-time set: any line length 2..4096 (factors 2, 3, 5, 7,
// any other prime as a direct DFT pass), shared with the C-ABI layer.
#pragma once
#include "exmi_internal.hpp"
#include "fft1d.hpp"

namespace exmi {

constexpr int GFFT_MAX_N = 4096;
constexpr int GFFT_MAX_PASS = 16;

// How a line length is transformed: the radices of the Stockham passes in
// order (their product is n) and how many lines one 256-thread workgroup
// holds in LDS. Chosen per size by the plan's autotuner (max_radix and the
// points per workgroup are the tuned knobs).
struct GfftCfg {
  int n = 0;
  int npass = 0;
  int radix[GFFT_MAX_PASS] = {};
  int lpb = 1;  // lines per workgroup
  int reg19 = 0;  // 1: radices 9..17 and 19 as register butterflies (a kernel of lower occupancy)
};

// Transforms of `nlines` lines.
//  C2C (forward or inverse, in place): line l = (outer o, inner i),
//      o = l / inner, i = l % inner, starts at o * outer_stride + i and its
//      points are pstride apart; the result is multiplied by `scale`.
//  R2C: real lines of n contiguous floats -> n/2+1 complex (contiguous).
//  C2R: n/2+1 complex (Hermitian half, Im of the DC and, for even n, the
//      Nyquist bin ignored) -> n real, multiplied by `scale`.
enum GfftMode { GFFT_FWD = 0, GFFT_INV = 1, GFFT_R2C = 2, GFFT_C2R = 3 };

// Radices for n with passes of at most max_radix points where n allows
// (prime factors above it form their own direct-DFT pass); lpb from the
// target number of points per workgroup. Returns false when n is outside
// 2..GFFT_MAX_N.
bool gfft_make_cfg(int n, int max_radix, int points_per_wg, GfftCfg* cfg);
// Encoded knobs (reg19 * 100000000 + max_radix * 100000 + points_per_wg) of
// the candidates the autotuner times, and the default used without tuning.
int gfft_num_candidates();
int gfft_candidate(int i);
int gfft_default_code();
bool gfft_cfg_from_code(int n, int code, GfftCfg* cfg);

hipError_t launch_gfft(const GfftCfg& c, int mode, const void* in, void* out, long long nlines, long long inner,
                       long long outer_stride, long long pstride, float scale, const cf* tw, hipStream_t st);

}  // namespace exmi
