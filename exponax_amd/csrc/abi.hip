// C-ABI layer: plan construction, dispatch to the fused kernels, hipGraph
// capture of rollouts, error reporting. See include/exmi.h for the contract.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <initializer_list>
#include <vector>

#include <climits>

#include "exmi_internal.hpp"
#include "adjoint1d.hpp"
#include "adjoint2d.hpp"
#include "fft1d.hpp"
#include "gfft.hpp"
#include "spectral3d.hpp"
#include "step1d.hpp"
#include "coop2d.hpp"
#include "roll2d.hpp"
#include "step2d.hpp"

namespace exmi {

using StepFn = hipError_t (*)(const Step1DArgs&, hipStream_t);
using RollFn = hipError_t (*)(const Roll1DArgs&, hipStream_t);
using VjpFn = hipError_t (*)(const Vjp1DArgs&, hipStream_t);
hipError_t launch_coef(const double* L, const double* Li, long long total, int nr, double dt, int order, int ncirc,
                       double radius, const double* nf, float* out, hipStream_t st);

hipError_t launch_lines(int N, bool inv, cf* data, long long nlines, long long inner, long long outer_stride,
                        long long pstride, float scale, const cf* tw, hipStream_t st, const cf* src = nullptr,
                        int kkeep = 0);
hipError_t launch_etdrk0_spec(cf* X, long long ninst, long long nr, int channels, const float* tables, int ncoef,
                              int lin_batch, int lin_channels, int e_im, hipStream_t st);
hipError_t launch_fields3d(const NL3DArgs& a, const cf* S, cf* F, hipStream_t st);
hipError_t launch_fields_lines3d(const NL3DArgs& a, const cf* S, cf* F, float scale, const cf* tw, int kkeep,
                                 hipStream_t st);
hipError_t launch_zcross3d(int N, const cf* F, cf* P, long long ntraj, long long Nr, int kcut, const RealTables& rt,
                           hipStream_t st);
hipError_t launch_zprod3d(const NL3DArgs& a, const cf* F, cf* P, int kcut, const RealTables& rt, hipStream_t st);
hipError_t launch_wave_spec(cf* X, long long batch, int D, int N, double kscale, double c, double dt,
                            hipStream_t st);
hipError_t launch_poisson_spec(cf* X, long long count, int D, int N, double kscale, int order, hipStream_t st);
hipError_t launch_axpy(long long n, float a, const float* x, const float* y, float* out, hipStream_t st);
hipError_t launch_products3d(const NL3DArgs& a, const float* R, float* PR, hipStream_t st);
hipError_t launch_combine3d(const NL3DArgs& a, const cf* PH, cf* Nh, hipStream_t st);
hipError_t launch_stage3d(const Stage3DArgs& a, hipStream_t st);
hipError_t launch_stage_nl3d(const NL3DArgs& na, const Stage3DArgs& a, const cf* PH, cf* N0out, hipStream_t st);
hipError_t launch_xstage_proj3d(const NL3DArgs& na, const Stage3DArgs& a, const cf* P, cf* N0out, int kcut,
                                const cf* tw, hipStream_t st, cf* K);
hipError_t launch_xstage3d(const NL3DArgs& na, const Stage3DArgs& a, const cf* P, cf* N0out, int kcut, const cf* tw,
                           hipStream_t st, cf* K);
hipError_t launch_fields_unit_lines3d(const NL3DArgs& a, const cf* S, cf* F, float scale, const cf* tw, int kkeep,
                                      hipStream_t st);
hipError_t launch_yz_inverse3d(int N, const cf* X, float* x, long long nplanes, const cf* tw, const RealTables& rt,
                               long long rpg, long long gstride, hipStream_t st);
hipError_t launch_yzprod3d(const NL3DArgs& a, const cf* F, cf* P, int kcut, const cf* tw, const RealTables& rt,
                           hipStream_t st);
hipError_t launch_adj_cofields3d(const NL3DArgs& a, const cf* MU, cf* H, hipStream_t st);
hipError_t launch_adj_products3d(const NL3DArgs& a, const float* R, const float* HR, float* Q, hipStream_t st);
hipError_t launch_adj_combine3d(const NL3DArgs& a, const cf* QH, cf* out, hipStream_t st);
hipError_t launch_adj_linear3d(const AdjLinArgs& a, hipStream_t st);

#define EXMI_DECL_SIZE(M)                                                                              \
  StepFn get_step1d_##M(int kind, int order);                                                          \
  RollFn get_roll1d_##M(int kind, int order, int nt, int fr, int std_kc);                              \
  StepFn get_step1d_c_##M(int kind, int order);                                                        \
  RollFn get_roll1d_c_##M(int kind, int order, int fr);                                                \
  RollFn get_roll1d_r_##M(int kind, int order);                                                        \
  VjpFn get_vjp1d_##M(int kind, int order, int cplx);                                                            \
  hipError_t launch_fft1d_##M(bool inverse, const float* in, float* out, int count, const RealTables& rt, \
                              hipStream_t st, int kcut, long long rpg, long long gstride);
EXMI_DECL_SIZE(8)
EXMI_DECL_SIZE(24)
EXMI_DECL_SIZE(25)
EXMI_DECL_SIZE(16)
EXMI_DECL_SIZE(32)
EXMI_DECL_SIZE(64)
EXMI_DECL_SIZE(128)
EXMI_DECL_SIZE(256)
EXMI_DECL_SIZE(512)
EXMI_DECL_SIZE(1024)
EXMI_DECL_SIZE(2048)

using Launch2D = hipError_t (*)(const Step2DArgs&, int, hipStream_t);
Launch2D get_2d_32(int kind, int is_row, int mode);
Launch2D get_2d_64(int kind, int is_row, int mode);
Launch2D get_2d_128(int kind, int is_row, int mode);
Launch2D get_2d_256(int kind, int is_row, int mode);
Launch2D get_2d_512(int kind, int is_row, int mode);
using LaunchRoll2D = hipError_t (*)(const Roll2DArgs&, hipStream_t);
LaunchRoll2D get_roll2d_32(int kind);
LaunchRoll2D get_roll2d_64(int kind);
LaunchRoll2D get_roll2d_128(int kind);
using LaunchCoop2D = hipError_t (*)(const Coop2DArgs&, hipStream_t);
using Coop2DCapacity = long long (*)();
LaunchCoop2D get_coop2d_32(int kind, Coop2DCapacity* cap);
LaunchCoop2D get_coop2d_64(int kind, Coop2DCapacity* cap);
LaunchCoop2D get_coop2d_128(int kind, Coop2DCapacity* cap);
LaunchCoop2D get_coop2d_256(int kind, Coop2DCapacity* cap);
LaunchCoop2D get_coop2d_512(int kind, Coop2DCapacity* cap);
using LaunchAdj2D = hipError_t (*)(const Adj2DArgs&, int, hipStream_t);
LaunchAdj2D get_adj2d_32(int is_row, int mode);
LaunchAdj2D get_adj2d_64(int is_row, int mode);
LaunchAdj2D get_adj2d_128(int is_row, int mode);
LaunchAdj2D get_adj2d_256(int is_row, int mode);
LaunchAdj2D get_adj2d_512(int is_row, int mode);

static LaunchRoll2D get_roll2d(int N, int kind) {
  switch (N) {
    case 32: return get_roll2d_32(kind);
    case 64: return get_roll2d_64(kind);
    case 128: return get_roll2d_128(kind);
    default: return nullptr;
  }
}

using LaunchCoopAdj2D = hipError_t (*)(const CoopAdj2DArgs&, hipStream_t);
LaunchCoopAdj2D get_coop_adj2d_32(Coop2DCapacity* cap);
LaunchCoopAdj2D get_coop_adj2d_64(Coop2DCapacity* cap);
LaunchCoopAdj2D get_coop_adj2d_128(Coop2DCapacity* cap);
LaunchCoopAdj2D get_coop_adj2d_256(Coop2DCapacity* cap);
LaunchCoopAdj2D get_coop_adj2d_512(Coop2DCapacity* cap);

static LaunchCoopAdj2D get_coop_adj2d(int N, Coop2DCapacity* cap) {
  *cap = nullptr;
  switch (N) {
    case 128: return get_coop_adj2d_128(cap);
    case 256: return get_coop_adj2d_256(cap);
    default: return nullptr;
  }
}

static LaunchCoop2D get_coop2d(int N, int kind, Coop2DCapacity* cap) {
  *cap = nullptr;
  switch (N) {
    case 128: return get_coop2d_128(kind, cap);
    case 256: return get_coop2d_256(kind, cap);
    default: return nullptr;
  }
}

static LaunchAdj2D get_adj2d(int N, int is_row, int mode) {
  switch (N) {
    case 32: return get_adj2d_32(is_row, mode);
    case 64: return get_adj2d_64(is_row, mode);
    case 128: return get_adj2d_128(is_row, mode);
    case 256: return get_adj2d_256(is_row, mode);
    case 512: return get_adj2d_512(is_row, mode);
    default: return nullptr;
  }
}

static Launch2D get_2d(int N, int kind, int is_row, int mode) {
  switch (N) {
    case 32: return get_2d_32(kind, is_row, mode);
    case 64: return get_2d_64(kind, is_row, mode);
    case 128: return get_2d_128(kind, is_row, mode);
    case 256: return get_2d_256(kind, is_row, mode);
    case 512: return get_2d_512(kind, is_row, mode);
    default: return nullptr;
  }
}

static StepFn get_step1d(int M, int kind, int order) {
  switch (M) {
    case 8: return get_step1d_8(kind, order);
    case 24: return get_step1d_24(kind, order);
    case 25: return get_step1d_25(kind, order);
    case 16: return get_step1d_16(kind, order);
    case 32: return get_step1d_32(kind, order);
    case 64: return get_step1d_64(kind, order);
    case 128: return get_step1d_128(kind, order);
    case 256: return get_step1d_256(kind, order);
    case 512: return get_step1d_512(kind, order);
    case 1024: return get_step1d_1024(kind, order);
    case 2048: return get_step1d_2048(kind, order);
    default: return nullptr;
  }
}

static VjpFn get_vjp1d(int M, int kind, int order, int cplx) {
  switch (M) {
    case 8: return get_vjp1d_8(kind, order, cplx);
    case 24: return get_vjp1d_24(kind, order, cplx);
    case 25: return get_vjp1d_25(kind, order, cplx);
    case 16: return get_vjp1d_16(kind, order, cplx);
    case 32: return get_vjp1d_32(kind, order, cplx);
    case 64: return get_vjp1d_64(kind, order, cplx);
    case 128: return get_vjp1d_128(kind, order, cplx);
    case 256: return get_vjp1d_256(kind, order, cplx);
    case 512: return get_vjp1d_512(kind, order, cplx);
    case 1024: return get_vjp1d_1024(kind, order, cplx);
    case 2048: return get_vjp1d_2048(kind, order, cplx);
    default: return nullptr;
  }
}

// Kernels for a complex linear operator (1-D, ETDRK orders 0-2, N <= 1024).
static StepFn get_step1d_c(int M, int kind, int order) {
  switch (M) {
#define EXMI_CASE(MM) \
  case MM: return get_step1d_c_##MM(kind, order);
    EXMI_CASE(8) EXMI_CASE(16) EXMI_CASE(24) EXMI_CASE(25) EXMI_CASE(32) EXMI_CASE(64) EXMI_CASE(128)
    EXMI_CASE(256) EXMI_CASE(512)
#undef EXMI_CASE
    default: return nullptr;
  }
}
static RollFn get_roll1d_c(int M, int kind, int order, int fr) {
  switch (M) {
#define EXMI_CASE(MM) \
  case MM: return get_roll1d_c_##MM(kind, order, fr);
    EXMI_CASE(8) EXMI_CASE(16) EXMI_CASE(24) EXMI_CASE(25) EXMI_CASE(32) EXMI_CASE(64) EXMI_CASE(128)
    EXMI_CASE(256) EXMI_CASE(512)
#undef EXMI_CASE
    default: return nullptr;
  }
}

// channel-coupled reactions: the channels of a trajectory share a lane group
static RollFn get_roll1d_r(int M, int kind, int order) {
  switch (M) {
#define EXMI_CASE(MM) \
  case MM: return get_roll1d_r_##MM(kind, order);
    EXMI_CASE(8) EXMI_CASE(16) EXMI_CASE(24) EXMI_CASE(25) EXMI_CASE(32) EXMI_CASE(64) EXMI_CASE(128)
    EXMI_CASE(256) EXMI_CASE(512)
#undef EXMI_CASE
    default: return nullptr;
  }
}

static RollFn get_roll1d(int M, int kind, int order, int nt, int fr, int std_kc) {
  switch (M) {
    case 8: return get_roll1d_8(kind, order, nt, fr, std_kc);
    case 24: return get_roll1d_24(kind, order, nt, fr, std_kc);
    case 25: return get_roll1d_25(kind, order, nt, fr, std_kc);
    case 16: return get_roll1d_16(kind, order, nt, fr, std_kc);
    case 32: return get_roll1d_32(kind, order, nt, fr, std_kc);
    case 64: return get_roll1d_64(kind, order, nt, fr, std_kc);
    case 128: return get_roll1d_128(kind, order, nt, fr, std_kc);
    case 256: return get_roll1d_256(kind, order, nt, fr, std_kc);
    case 512: return get_roll1d_512(kind, order, nt, fr, std_kc);
    case 1024: return get_roll1d_1024(kind, order, nt, fr, std_kc);
    case 2048: return get_roll1d_2048(kind, order, nt, fr, std_kc);
    default: return nullptr;
  }
}

static hipError_t launch_fft1d(int M, bool inverse, const float* in, float* out, int count, const RealTables& rt,
                               hipStream_t st, int kcut = 1 << 30, long long rpg = 0, long long gstride = 0) {
  switch (M) {
#define EXMI_CASE(MM) \
  case MM: return launch_fft1d_##MM(inverse, in, out, count, rt, st, kcut, rpg, gstride);
    EXMI_CASE(8)
    EXMI_CASE(16)
    EXMI_CASE(24)
    EXMI_CASE(25)
    EXMI_CASE(32)
    EXMI_CASE(64)
    EXMI_CASE(128)
    EXMI_CASE(256)
    EXMI_CASE(512)
    EXMI_CASE(1024)
    EXMI_CASE(2048)
#undef EXMI_CASE
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// the error channel for the library's other translation units
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(EXMI_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------- twiddle table cache
struct Twiddles {
  cf* tw = nullptr;  // exp(-2 pi i m / M), m < M
  cf* fw = nullptr;  // -i w_k, w_k = exp(-2 pi i k / (2M)), k < M
  RealTables rt() const { return RealTables{tw, fw}; }
};

static std::mutex g_tw_mu;
static std::map<std::pair<int, int>, Twiddles> g_tw;  // (device, M)

// Host values of the two tables of transform size M, formed in double
// precision and rounded once to float.
static void twiddle_values(int M, cf* h, cf* hf) {
  const double pi = 3.14159265358979323846;
  for (int m = 0; m < M; ++m) h[m] = cf{(float)std::cos(2 * pi * m / M), (float)-std::sin(2 * pi * m / M)};
  for (int k = 0; k < M; ++k) {
    const double c = std::cos(pi * k / M), s = -std::sin(pi * k / M);  // w_k
    hf[k] = cf{(float)s, (float)(-c)};                                  // -i w_k
  }
}

static int get_twiddles(int M, Twiddles* out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, M);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) {
    *out = it->second;
    return EXMI_OK;
  }
  std::vector<cf> h(M), hf(M);
  twiddle_values(M, h.data(), hf.data());
  Twiddles t;
  if (hipMalloc(&t.tw, sizeof(cf) * M) != hipSuccess || hipMalloc(&t.fw, sizeof(cf) * M) != hipSuccess)
    return fail(EXMI_ERR_ALLOC, "twiddle allocation failed");
  HIP_TRY(hipMemcpy(t.tw, h.data(), sizeof(cf) * M, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.fw, hf.data(), sizeof(cf) * M, hipMemcpyHostToDevice));
  g_tw[key] = t;
  *out = t;
  return EXMI_OK;
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
// 1-D transform sizes: powers of two 16..4096, and 48 and 50 (single-lane
// mixed-radix transforms of 24 = 8 x 3 and 25 = 5 x 5 points)
static bool size_1d(int n) { return (pow2(n) && n >= 16 && n <= 4096) || n == 48 || n == 50; }

// Dealias cutoff of nonlin_fun/_base.py:59-71: keep |k| <= frac*(N//2) - 1,
// compared in float32 as the reference's float32 wavenumber grid is.
// Per-row factors of the nonlinear-term coefficients: row r of the tables
// (lin_batch, lin_channels) takes host_nonlin_factor[r / lin_channels];
// nullptr without factors. The caller frees *out.
static int upload_nonlin_factor(const exmi_plan_desc* d, long long rows, double** out) {
  *out = nullptr;
  if (!d->host_nonlin_factor) return EXMI_OK;
  std::vector<double> h((size_t)rows);
  for (long long r = 0; r < rows; ++r) h[(size_t)r] = d->host_nonlin_factor[r / d->lin_channels];
  HIP_TRY(hipMalloc(out, sizeof(double) * rows));
  const hipError_t e = hipMemcpy(*out, h.data(), sizeof(double) * rows, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(*out);
    *out = nullptr;
    return fail(EXMI_ERR_HIP, "factor upload failed: %s", hipGetErrorString(e));
  }
  return EXMI_OK;
}

static int dealias_kmax(int N, double frac) {
  if (frac < 0) return 1 << 30;
  const double cutoff = frac * (double)(N / 2) - 1.0;
  const float cf32 = (float)cutoff;
  if (cf32 < 0.f) return -1;
  return (int)std::floor((double)cf32);
}

}  // namespace exmi

using namespace exmi;

struct GraphCache {
  const void* a = nullptr;
  const void* b = nullptr;
  long long s = 0;
  int n = -1, flag = -1, kind = -1;
  hipGraphExec_t exec = nullptr;
};

struct exmi_plan {
  exmi_plan_desc d{};
  int M = 0, nr = 0, ncoef = 0, kind = 0;
  int kmax = 0;
  float kscale = 0.f;
  int device = 0;
  float* tables = nullptr;
  float* work = nullptr;
  size_t work_elems = 0;
  Twiddles tw;
  StepFn step = nullptr;
  StepFn nonlin = nullptr;
  RollFn roll[3] = {nullptr, nullptr, nullptr};  // by transforms per lane group (1, 2)
  int roll_nt = 0;                               // chosen by the autotuner (0 = not yet)
  int tune_short = 0;  // chosen on a probe shorter than kTuneProbe steps: timed again on a longer rollout
  int fourier_resident = 0;                      // rollouts keep the spectral state on chip
  int fr_auto = 0;  // chosen by the plan (exact identity); per-step launches stay allowed
  hipStream_t capture_stream = nullptr;
  hipStream_t launch_stream = nullptr;  // concurrent group rollouts
  GraphCache gc;
  // 2-D plans
  int ndim = 1;
  int cplx = 0;  // 2-D: complex linear operator (Im E follows the real coefficients)
  int coupled = 0;  // 1-D channel-coupled reaction: steps through the multi-step kernel
  int nf = 0, kc = 0;
  int prod_nc = 0;  // 2-D/3-D general kind: first product u sum_d d_d u (single-channel non-conservative)
  int ninst = 0;          // 2-D and 3-D: instances = batch * channels
  long long tstride = 0;  // 2-D: complex entries of T per instance
  cf* W3 = nullptr;       // 3-D: spectra of every instance (the step's U0)
  cf* K3 = nullptr;       // multi-pass spectrum-carrying rollout: the emitted states' inverse transform
  int md_carry = 0;       // multi-pass plans: the spectrum-carrying rollout (strategy 2) available
  float* P3 = nullptr;    // 3-D: contiguous staging of strided physical states
  // 3-D nonlinear steps: stage spectra, nonlinear terms and evaluation buffers
  cf *N03 = nullptr, *NC3 = nullptr, *S3 = nullptr, *S13 = nullptr, *A3 = nullptr, *F3 = nullptr;
  float *R3 = nullptr, *PR3 = nullptr;
  int nf3 = 0, np3 = 0, kind3 = 0;
  // multi-pass spectral plans (fields / products / combine / stage passes
  // around whole-field transforms): 3-D plans, and plans of any dimension
  // whose grid size is outside the fused kernels' set
  int md = 0;
  int gen = 0;           // runtime-size transforms (gfft.hip)
  int gnx = 1, gny = 1;  // sizes of the first two of the three embedded axes
  long long n3 = 0;      // points of one field, N^D
  GfftCfg gcfg;          // runtime-size transforms: radices and lines per workgroup
  int gcode = 0;         // their knobs, encoded (autotuned at plan creation)
  int gtuned = 0;        // 0: default knobs, 1: timed at plan creation, 2: from the tuning cache, 3: set
  // multi-pass adjoint: stage states, cotangents and the nonlinear VJP's
  // cotangent fields / products (allocated on first use)
  cf *XA3 = nullptr, *XB3 = nullptr, *XC3 = nullptr, *LAM3 = nullptr, *MU3 = nullptr, *T3 = nullptr;
  cf *LA3 = nullptr, *LB3 = nullptr, *LC3 = nullptr, *FH3 = nullptr;
  float *RH3 = nullptr, *Q3 = nullptr;
  float* Gseg = nullptr;  // 2-D rollout gradient: one segment's gradient fields from the forward launch
  long long gseg_elems = 0;
  int adj_nh = 0, adj_nq = 0;
  // fused plans whose adjoint the fused kernels do not cover: the
  // description and host operator tables kept to build a multi-pass twin
  // plan of the same stepper on first use
  exmi_plan* twin = nullptr;
  exmi_plan_desc d0{};
  std::vector<double> hL, hLi, hF;
  Launch2D row2d[ROW_NUM_MODES] = {};
  LaunchRoll2D roll2d = nullptr;  // 2-D workgroup-resident rollout (strategy 1)
  int fr2d = 0;                   // 2-D spectrum-resident launch chain available (strategy 2)
  LaunchCoop2D coop2d = nullptr;  // the chain as one cooperative launch (strategy 3)
  LaunchCoopAdj2D coop_adj = nullptr;  // transposed rollout segments as one cooperative launch
  unsigned* coop_sync = nullptr;  // its arrival counters and failure word
  int coop_spin = 1 << 20;        // polls of a cooperative barrier before it gives up
  Launch2D col2d[COL_NUM_MODES] = {};
  Twiddles tw_col;
  cf *S0 = nullptr, *N0 = nullptr, *S1 = nullptr, *F = nullptr, *T2 = nullptr, *ACC = nullptr;
  float inj_value = 0.f;
  int inj_mode = -1;
  // adjoint: 1-D kernel; 2-D modes and workspaces (allocated on first use)
  VjpFn vjp = nullptr;
  LaunchAdj2D rowA[3] = {nullptr, nullptr, nullptr}, colA[3] = {nullptr, nullptr, nullptr};
  float* G = nullptr;
  cf *LAM = nullptr, *RHO = nullptr, *H = nullptr, *P = nullptr;
  float* ckpt = nullptr;  // rollout-gradient checkpoints, segment states, accumulator
  size_t ckpt_elems = 0;
};

// ---- grouped rollouts: several plans (e.g. different grid sizes) advanced
// together; in a graph each plan's launches form a parallel branch.
struct GroupGraph {
  std::vector<const void*> key;
  int n = -1, flag = -1;
  hipGraphExec_t exec = nullptr;
};
static std::mutex g_group_mu;
static std::map<const exmi_plan*, GroupGraph> g_group_graphs;  // by leader plan

static void fill_args2d(const exmi_plan* p, Step2DArgs* a) {
  memset(a, 0, sizeof *a);
  a->S0 = p->S0;
  a->N0 = p->N0;
  a->S1 = p->S1;
  a->F = p->F;
  a->T = p->T2;
  a->A = p->ACC;
  a->tables = p->tables;
  a->ncoef = p->ncoef;
  a->lin_batch = p->d.lin_batch;
  a->tw_row = p->tw.tw;
  a->fw_row = p->tw.fw;
  a->tw_col = p->tw_col.tw;
  a->kscale = p->kscale;
  a->kmax = p->kmax;
  a->scale = (float)p->d.scale;
  a->zero_mode_fix = p->d.zero_mode_fix;
  a->inj_mode = p->inj_mode;
  a->inj_value = p->inj_value;
  a->channels = p->d.num_channels;
  a->lin_channels = p->d.lin_channels;
  a->feed = (float)p->d.reaction_rates[0];
  a->kill = (float)p->d.reaction_rates[1];
  a->tstride = p->tstride;
  a->g0 = (float)p->d.general_scales[0];
  a->g1 = (float)p->d.general_scales[1];
  a->g2 = (float)p->d.general_scales[2];
  a->prod_nc = p->prod_nc;
  a->e_im = p->cplx ? num_coefs(p->d.order) : -1;
  a->poly_lap = p->d.nonlin_kind == EXMI_NL_CAHN_HILLIARD;
  a->npoly = p->d.num_poly;
  for (int i = 0; i < EXMI_MAX_POLY; ++i) a->poly[i] = (float)p->d.poly[i];
}

static void fill_args(const exmi_plan* p, Step1DArgs* a) {
  memset(a, 0, sizeof *a);
  a->ntrans = p->d.batch * p->d.num_channels;
  a->channels = p->d.num_channels;
  a->tables = p->tables;
  a->lin_batch = p->d.lin_batch;
  a->lin_channels = p->d.lin_channels;
  a->ncoef = p->ncoef;
  a->tw = p->tw.tw;
  a->fw = p->tw.fw;
  a->kscale = p->kscale;
  a->kmax = p->kmax;
  a->scale = (float)p->d.scale;
  a->g0 = (float)p->d.general_scales[0];
  a->g1 = (float)p->d.general_scales[1];
  a->g2 = (float)p->d.general_scales[2];
  a->zero_mode_fix = p->d.zero_mode_fix;
  a->npoly = p->d.num_poly;
  for (int i = 0; i < EXMI_MAX_POLY; ++i) a->poly[i] = (float)p->d.poly[i];
  a->poly_lap = p->d.nonlin_kind == EXMI_NL_CAHN_HILLIARD;
  a->poly_lap_c = (float)(-p->d.scale * (double)p->kscale * (double)p->kscale);
  a->feed = (float)p->d.reaction_rates[0];
  a->kill = (float)p->d.reaction_rates[1];
}

// Cahn-Hilliard's u^3 as the polynomial (0, 0, 0, 1); the kernels apply
// scale * Laplace to its transform.
static void cahn_hilliard_poly(exmi_plan_desc* d) {
  d->num_poly = 4;
  for (int i = 0; i < EXMI_MAX_POLY; ++i) d->poly[i] = i == 3 ? 1.0 : 0.0;
}

static int kernel_kind(const exmi_plan_desc* d, int* kind) {
  const int D = d->num_spatial_dims, C = d->num_channels;
  switch (d->nonlin_kind) {
    case EXMI_NL_ZERO: *kind = NL_ZERO; return EXMI_OK;
    case EXMI_NL_CONVECTION:
      if (d->conv_single_channel) {
        if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "single-channel convection needs num_channels == 1");
      } else if (C != D) {
        // nonlin_fun/_convection.py:125-129
        return fail(EXMI_ERR_INVALID, "Number of channels in u_hat should match number of spatial dimensions");
      }
      *kind = d->conv_conservative ? NL_CONV_CONS : NL_CONV_NONCONS;
      return EXMI_OK;
    case EXMI_NL_GRADIENT_NORM: *kind = NL_GRADNORM; return EXMI_OK;
    case EXMI_NL_POLYNOMIAL:
      if (d->num_poly < 0 || d->num_poly > EXMI_MAX_POLY)
        return fail(EXMI_ERR_UNSUPPORTED, "at most %d polynomial coefficients", EXMI_MAX_POLY);
      *kind = NL_POLY;
      return EXMI_OK;
    case EXMI_NL_CAHN_HILLIARD:
      if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "Cahn-Hilliard needs num_channels == 1");
      *kind = NL_POLY;
      return EXMI_OK;
    case EXMI_NL_GRAY_SCOTT:  // _gray_scott.py:35-36
      if (C != 2) return fail(EXMI_ERR_INVALID, "num_channels must be 2");
      *kind = NL_GRAY_SCOTT;
      return EXMI_OK;
    case EXMI_NL_BELOUSOV_ZHABOTINSKY:  // _belousov_zhabotinsky.py:35-37
      if (C != 3) return fail(EXMI_ERR_INVALID, "num_channels must be 3");
      *kind = NL_BZ;
      return EXMI_OK;
    case EXMI_NL_GENERAL:
      if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "general nonlinear function needs num_channels == 1");
      *kind = NL_GENERAL;
      return EXMI_OK;
    case EXMI_NL_VORTICITY:
    case EXMI_NL_KOLMOGOROV:
      if (D != 2) return fail(EXMI_ERR_INVALID, "Expected num_spatial_dims = 2, got %d.", D);
      *kind = NL_VORT;
      return EXMI_OK;
    default: return fail(EXMI_ERR_INVALID, "unknown nonlinear kind %d", d->nonlin_kind);
  }
}

int exmi_plan_destroy(exmi_plan* p);
static int ensure_work(exmi_plan* p);
static void keep_for_twin(exmi_plan* p, const exmi_plan_desc* d, long long total);

// 2-D plans: ETDRK0-4 with the vorticity (Kolmogorov), gradient-norm,
// polynomial, two-channel convection or general nonlinearity (or none), N a
// power of two in 32..512; complex linear operators for orders 0-2.
static int plan_create_2d(const exmi_plan_desc* d, exmi_plan** out) {
  const int N = d->num_points;
  const int C = d->num_channels;
  if (!pow2(N) || N < 32 || N > 512)
    return fail(EXMI_ERR_UNSUPPORTED, "2-D num_points=%d: supported sizes are powers of two 32..512", N);
  if (C < 1) return fail(EXMI_ERR_INVALID, "num_channels must be >= 1");
  if (d->batch < 1) return fail(EXMI_ERR_INVALID, "batch must be >= 1");
  if (d->order < 0 || d->order > 4) return fail(EXMI_ERR_UNSUPPORTED, "Order %d not implemented.", d->order);
  if (d->flags & EXMI_FLAG_FOURIER_RESIDENT)
    return fail(EXMI_ERR_UNSUPPORTED, "Fourier-resident rollouts are built for 1-D plans");
  if (d->lin_batch != 1 && d->lin_batch != d->batch)
    return fail(EXMI_ERR_INVALID, "lin_batch must be 1 or batch (%d), got %d", d->batch, d->lin_batch);
  if (d->lin_channels != 1 && d->lin_channels != C)
    return fail(EXMI_ERR_INVALID, "lin_channels must be 1 or num_channels (%d), got %d", C, d->lin_channels);
  if (!d->host_linear_operator) return fail(EXMI_ERR_INVALID, "host_linear_operator is null");
  const bool cplx = d->host_linear_operator_imag != nullptr;
  int kind = 0;
  bool zero_poly = false, conv_general = false;
  if (d->nonlin_kind == EXMI_NL_VORTICITY || d->nonlin_kind == EXMI_NL_KOLMOGOROV) {
    if (C != 1) return fail(EXMI_ERR_INVALID, "vorticity plans need num_channels == 1");
    kind = NL_VORT;
  } else if (d->nonlin_kind == EXMI_NL_CONVECTION && d->order > 0) {
    if (d->conv_single_channel) {
      if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "single-channel convection needs num_channels == 1");
      // -b/2 (sum_d i kappa_d) F[u^2] is the general form with scales
      // (0, -b, 0); the non-conservative -b F[u sum_d d_d u] is the general
      // form's fields with that product and scales (-b, 0, 0)
      kind = NL_GENERAL;
      conv_general = true;
    } else {
      if (C != 2)  // nonlin_fun/_convection.py:125-129
        return fail(EXMI_ERR_INVALID, "Number of channels in u_hat should match number of spatial dimensions");
      kind = d->conv_conservative ? NL_CONV_CONS : NL_CONV_NONCONS;
    }
  } else if (d->nonlin_kind == EXMI_NL_GENERAL && d->order > 0) {
    if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "general nonlinear function needs num_channels == 1");
    kind = NL_GENERAL;
  } else if (d->nonlin_kind == EXMI_NL_GRADIENT_NORM || d->order == 0) {
    kind = NL_GRADNORM;  // ETDRK0 evaluates no nonlinear term
  } else if (d->nonlin_kind == EXMI_NL_POLYNOMIAL) {
    if (d->num_poly < 0 || d->num_poly > EXMI_MAX_POLY)
      return fail(EXMI_ERR_UNSUPPORTED, "at most %d polynomial coefficients", EXMI_MAX_POLY);
    kind = NL_POLY;
  } else if (d->nonlin_kind == EXMI_NL_CAHN_HILLIARD) {
    if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "Cahn-Hilliard needs num_channels == 1");
    kind = NL_POLY;
  } else if (d->nonlin_kind == EXMI_NL_GRAY_SCOTT) {
    if (C != 2) return fail(EXMI_ERR_INVALID, "num_channels must be 2");
    kind = NL_GRAY_SCOTT;
  } else if (d->nonlin_kind == EXMI_NL_BELOUSOV_ZHABOTINSKY) {
    if (C != 3) return fail(EXMI_ERR_INVALID, "num_channels must be 3");
    kind = NL_BZ;
  } else if (d->nonlin_kind == EXMI_NL_ZERO) {
    kind = NL_POLY;  // the empty polynomial: N = 0 exactly
    zero_poly = true;
  } else {
    return fail(EXMI_ERR_UNSUPPORTED, "unknown 2-D nonlinear kind %d", d->nonlin_kind);
  }
  exmi_plan* p = new exmi_plan();
  p->d = *d;
  p->d.host_linear_operator = nullptr;
  p->d.host_linear_operator_imag = nullptr;
  p->ndim = 2;
  p->M = N / 2;
  const int NY = N / 2 + 1;
  p->nr = N * NY;
  p->ncoef = num_coefs_c(d->order, cplx);
  p->cplx = cplx;
  p->kind = kind;
  if (zero_poly) p->d.num_poly = 0;
  if (d->nonlin_kind == EXMI_NL_CAHN_HILLIARD) cahn_hilliard_poly(&p->d);
  if (conv_general) {
    p->prod_nc = d->conv_conservative ? 0 : 1;
    p->d.general_scales[0] = p->prod_nc ? -d->scale : 0.0;
    p->d.general_scales[1] = p->prod_nc ? 0.0 : -d->scale;
    p->d.general_scales[2] = 0.0;
  }
  p->kmax = dealias_kmax(N, d->dealiasing_fraction);
  p->kc = p->kmax + 1 < NY ? p->kmax + 1 : NY;
  if (p->kc < 1) p->kc = 1;
  {
    const int np = (kind == NL_CONV_CONS || kind == NL_GENERAL) ? 2 : 1;
    p->nf = kind == NL_VORT ? 4
            : kind == NL_GRADNORM ? 2
            : (kind == NL_POLY || kind == NL_CONV_CONS || kind == NL_GRAY_SCOTT || kind == NL_BZ) ? 1
                                                                                                  : 3;
    const long long fullc = (long long)NY * N, prod = (long long)np * p->kc * N;
    p->tstride = fullc > prod ? fullc : prod;
    p->ninst = d->batch * C;
  }
  p->kscale = (float)(2.0 * 3.14159265358979323846 / d->domain_extent);
  p->roll_nt = -1;
  hipGetDevice(&p->device);
  // Workgroup-resident ETDRK2 rollouts (one workgroup keeps an instance's
  // spectrum on chip for all steps): N <= 128, single-channel gradient norm
  // or vorticity, real L, dealias cutoff up to the default 2/3 one (the
  // Nyquist row and column then never reach the nonlinear term, so the
  // per-step forward transform of the emitted state is an exact identity),
  // forcing below the Nyquist column. Other plans, and every plan with
  // EXMI_FLAG_PHYSICAL_ROLLOUT, step through the row/column launches.
  if (d->order == 2 && !cplx && C == 1 && N <= 128 && !(d->flags & EXMI_FLAG_PHYSICAL_ROLLOUT) &&
      (d->nonlin_kind == EXMI_NL_GRADIENT_NORM || kind == NL_VORT) && p->kmax >= 0 &&
      p->kmax <= (2 * (N / 2) - 3) / 3 && d->injection_mode < N / 2)
    p->roll2d = get_roll2d(N, kind);
  // The same exactness lets the row/column launches carry S' in the column
  // layout from step to step (ETDRK2, every kind: the products of real
  // fields have Hermitian transforms): four launches per step instead of
  // six (no forward transform of the emitted state).
  p->fr2d = d->order == 2 && !cplx && p->kmax >= 0 && p->kmax < N / 2 && !(d->flags & EXMI_FLAG_PHYSICAL_ROLLOUT);
  // The chain as one launch whose workgroups meet at barriers between the
  // passes (single-channel gradient norm and vorticity, N = 128 and 256),
  // where the whole grid fits the device at once.
  if (p->fr2d && C == 1) {
    // EXMI_COOP_SPIN_LIMIT: the barriers' poll limit (tests force the
    // give-up path with a tiny one; the launch then ends with NaN states)
    if (const char* sl = getenv("EXMI_COOP_SPIN_LIMIT")) p->coop_spin = atoi(sl) > 0 ? atoi(sl) : 1;
    Coop2DCapacity cap = nullptr;
    LaunchCoop2D f = get_coop2d(N, kind, &cap);
    if (f && cap && p->ninst <= cap() &&
        hipMalloc(&p->coop_sync, sizeof(unsigned) * (64 * (size_t)p->ninst + 64)) == hipSuccess)
      p->coop2d = f;
    (void)hipGetLastError();
  }
  if (p->roll2d || p->fr2d) p->roll_nt = 0;  // the autotuner picks among them and the per-step launches
  if (d->nonlin_kind == EXMI_NL_KOLMOGOROV) {
    // _vorticity_convection.py:166-176: -k_f gamma scaling(coef_extraction)
    // at wavenumber (0, k_f); scaling = N at k = 0 or N/2, else N/2 per axis
    const int kf = d->injection_mode;
    if (kf >= 0 && kf <= N / 2) {
      const double sy = (kf == 0 || kf == N / 2) ? N : N / 2.0;
      p->inj_mode = kf;
      p->inj_value = (float)(-(double)kf * d->injection_scale * (double)N * sy);
    }
  }
  for (int m = 0; m < ROW_NUM_MODES; ++m) p->row2d[m] = get_2d(N, kind, 1, m);
  for (int m = 0; m < COL_NUM_MODES; ++m) p->col2d[m] = get_2d(N, kind, 0, m);
  for (int m = 0; m < COL_NUM_MODES; ++m) {
    if (m == COL_FWD || m == COL_INV) continue;  // the standalone transforms' modes
    if ((m < ROW_NUM_MODES && !p->row2d[m]) || !p->col2d[m]) {
      delete p;
      return fail(EXMI_ERR_UNSUPPORTED, "no 2-D kernel for N=%d", N);
    }
  }
  if (kind == NL_GRADNORM && d->order == 2 && !cplx && C == 1) {
    for (int m = 0; m < 3; ++m) {
      p->rowA[m] = get_adj2d(N, 1, m);
      p->colA[m] = get_adj2d(N, 0, m);
    }
    // with the cooperative forward, the reverse pass of a rollout gradient
    // runs each checkpoint segment as one cooperative launch too
    Coop2DCapacity cap = nullptr;
    LaunchCoopAdj2D f = p->coop2d ? get_coop_adj2d(N, &cap) : nullptr;
    if (f && cap && p->ninst <= cap()) p->coop_adj = f;
    (void)hipGetLastError();
  }
  int rc = get_twiddles(N / 2, &p->tw);
  if (!rc) rc = get_twiddles(N, &p->tw_col);
  if (rc) {
    delete p;
    return rc;
  }
  const long long B = p->ninst, full = (long long)NY * N;  // allocations per instance
  const long long rows = (long long)d->lin_batch * d->lin_channels;
  // L arrives as the reference's [kx][ky] table; the column kernels read
  // [ky][kx]
  // [ky][kx]; with a complex L the imaginary part follows the real one
  const int nparts = cplx ? 2 : 1;
  std::vector<double> Lt((size_t)(nparts * rows * full));
  for (int part = 0; part < nparts; ++part) {
    const double* src = part ? d->host_linear_operator_imag : d->host_linear_operator;
    double* dst = Lt.data() + (size_t)part * rows * full;
    for (long long b = 0; b < rows; ++b)
      for (int kx = 0; kx < N; ++kx)
        for (int ky = 0; ky < NY; ++ky)
          dst[(size_t)(b * full + (long long)ky * N + kx)] = src[b * full + (long long)kx * NY + ky];
  }
  double* dL = nullptr;
  bool ok = hipMalloc(&dL, sizeof(double) * nparts * rows * full) == hipSuccess &&
            hipMalloc(&p->tables, sizeof(float) * rows * full * p->ncoef) == hipSuccess &&
            hipMalloc(&p->S0, sizeof(cf) * B * full) == hipSuccess &&
            hipMalloc(&p->S1, sizeof(cf) * B * full) == hipSuccess &&
            hipMalloc(&p->T2, sizeof(cf) * B * p->tstride) == hipSuccess &&
            hipMalloc(&p->N0, sizeof(cf) * B * p->kc * N) == hipSuccess &&
            hipMalloc(&p->F, sizeof(cf) * B * p->nf * p->kc * N) == hipSuccess &&
            (d->order < 3 || hipMalloc(&p->ACC, sizeof(cf) * B * full) == hipSuccess);
  if (!ok) {
    hipFree(dL);
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_ALLOC, "2-D workspace allocation failed (batch %lld, N %d)", B, N);
  }
  double* dF = nullptr;
  int frc = upload_nonlin_factor(d, rows, &dF);
  if (frc) {
    hipFree(dL);
    exmi_plan_destroy(p);
    return frc;
  }
  hipError_t e = hipMemcpy(dL, Lt.data(), sizeof(double) * nparts * rows * full, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_coef(dL, cplx ? dL + rows * full : nullptr, rows * full, (int)full, d->dt, d->order,
                    d->num_circle_points, d->circle_radius, dF, p->tables, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  hipFree(dL);
  hipFree(dF);
  if (e != hipSuccess) {
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_HIP, "coefficient formation failed: %s", hipGetErrorString(e));
  }
  if (hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&p->launch_stream, hipStreamNonBlocking) != hipSuccess) {
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_HIP, "stream creation failed");
  }
  if (!p->rowA[0]) keep_for_twin(p, d, rows * full);  // the multi-pass adjoint, on first use
  *out = p;
  return EXMI_OK;
}

// One 2-D ETDRK2 step: u_in (stride ibs) -> u_out (stride obs). With
// have_T the row spectra of u_in are already in T (rollout); with write_T
// the last launch also leaves the new state's row spectra in T.
// floats of one state (C, N^D)
static long long state_size(const exmi_plan* p) {
  long long s = p->d.num_channels;
  for (int i = 0; i < p->ndim; ++i) s *= p->d.num_points;
  return s;
}

// ------------------------------------------------- multi-pass plans (3-D, any size)
static bool size_3d(int N) { return pow2(N) && N >= 16 && N <= 256; }
static bool size_2d(int N) { return pow2(N) && N >= 32 && N <= 512; }
// Sizes of the fused kernels; every other N in 2..GFFT_MAX_N runs through
// the multi-pass plans with the runtime-size transforms of gfft.hip.
static bool fused_size(int D, int N) { return D == 1 ? size_1d(N) : D == 2 ? size_2d(N) : size_3d(N); }
static bool gen_size(int N) { return N >= 2 && N <= GFFT_MAX_N; }
static hipError_t copy_batch(const float* src, long long ss, float* dst, long long ds, long long n, int B,
                             hipStream_t st);

// rfftn of I contiguous N^3 fields x into spectra X (I, N, N, N/2+1): z rows
// (r2c), then y lines and x lines (c2c) in place.
static hipError_t rfft3(int N, long long I, const float* x, cf* X, const Twiddles& tz, const Twiddles& tl,
                        hipStream_t st) {
  const long long NZ = N / 2 + 1;
  hipError_t e = launch_fft1d(N / 2, false, x, (float*)X, (int)(I * N * N), tz.rt(), st);
  if (e == hipSuccess) e = launch_lines(N, false, X, I * N * NZ, NZ, N * NZ, NZ, 1.f, tl.tw, st);
  if (e == hipSuccess) e = launch_lines(N, false, X, I * N * NZ, N * NZ, (long long)N * N * NZ, N * NZ, 1.f, tl.tw, st);
  return e;
}

// irfftn: the x and y inverse lines in place on X (each scaled by 1/N), then
// the z rows c2r (scaled by 1/N, Im of kz = 0 and N/2 ignored) into x. With
// src, the x lines read src instead and X is only the intermediate (src is
// left unchanged).
static hipError_t irfft3(int N, long long I, cf* X, float* x, const Twiddles& tz, const Twiddles& tl,
                         hipStream_t st, const cf* src = nullptr, long long rpg = 0, long long gstride = 0) {
  const long long NZ = N / 2 + 1;
  const float r = 1.f / (float)N;
  hipError_t e = launch_lines(N, true, X, I * N * NZ, N * NZ, (long long)N * N * NZ, N * NZ, r, tl.tw, st, src);
  // up to N = 64 the y lines and the z c2r run as one plane kernel
  if (N <= 64) return e == hipSuccess ? launch_yz_inverse3d(N, X, x, I * N, tl.tw, tz.rt(), rpg, gstride, st) : e;
  if (e == hipSuccess) e = launch_lines(N, true, X, I * N * NZ, NZ, N * NZ, NZ, r, tl.tw, st);
  if (e == hipSuccess)
    e = launch_fft1d(N / 2, true, (const float*)X, x, (int)(I * N * N), tz.rt(), st, 1 << 30, rpg, gstride);
  return e;
}

// rfftn over the last D axes of I contiguous N^D fields with the
// runtime-size transforms: r2c along the last axis, then c2c in place along
// the others (innermost first). tw: the N-th roots of unity.
static hipError_t rfft_gen(const GfftCfg& c, int D, long long I, const float* x, cf* X, const cf* tw,
                           hipStream_t st) {
  const int N = c.n;
  const long long NZ = N / 2 + 1;
  long long total = I * NZ;  // complex entries of all spectra
  for (int d = 1; d < D; ++d) total *= N;
  hipError_t e = launch_gfft(c, GFFT_R2C, x, X, total / NZ, 1, 0, 1, 1.f, tw, st);
  long long inner = NZ;  // entries between neighbouring points of the axis
  for (int ax = D - 2; ax >= 0 && e == hipSuccess; --ax) {
    e = launch_gfft(c, GFFT_FWD, X, X, total / N, inner, inner * N, inner, 1.f, tw, st);
    inner *= N;
  }
  return e;
}

// irfftn: the inverse c2c lines in place on X (outermost axis first, each
// scaled by 1/N), then c2r along the last axis into x. With src, the first
// pass reads src instead and X is only the intermediate (src is left
// unchanged; in 1-D the c2r reads src).
static hipError_t irfft_gen(const GfftCfg& c, int D, long long I, cf* X, float* x, const cf* tw, hipStream_t st,
                            const cf* src = nullptr) {
  const int N = c.n;
  const long long NZ = N / 2 + 1;
  const float r = 1.f / (float)N;
  long long total = I * NZ, inner = NZ;
  for (int d = 1; d < D; ++d) total *= N;
  for (int d = 0; d < D - 2; ++d) inner *= N;
  hipError_t e = hipSuccess;
  const cf* in = src ? src : X;
  for (int ax = 0; ax <= D - 2 && e == hipSuccess; ++ax) {
    e = launch_gfft(c, GFFT_INV, in, X, total / N, inner, inner * N, inner, r, tw, st);
    in = X;
    inner /= N;
  }
  if (e == hipSuccess) e = launch_gfft(c, GFFT_C2R, in, x, total / NZ, 1, 0, 1, r, tw, st);
  return e;
}

static hipError_t rfft_md(const exmi_plan* p, long long I, const float* x, cf* X, hipStream_t st);
static hipError_t irfft_md(const exmi_plan* p, long long I, cf* X, float* x, hipStream_t st);
static int tune_gfft(exmi_plan* p);

// Multi-pass plans: 3-D grids of the fused sizes (powers of two 16..256),
// and grids of any dimension whose size is outside the fused kernels' set
// (gen: runtime-size transforms, N in 2..GFFT_MAX_N). ETDRK0-4 (real or
// complex L) with the polynomial (Cahn-Hilliard), gradient-norm, convection
// (all four forms), general, vorticity (2-D), projected-convection (3-D) or
// coupled-reaction nonlinearity (or none).
static int plan_create_md(const exmi_plan_desc* d, exmi_plan** out, bool gen) {
  const int N = d->num_points, C = d->num_channels, D = d->num_spatial_dims;
  if (!gen && !size_3d(N))
    return fail(EXMI_ERR_UNSUPPORTED, "3-D num_points=%d: supported sizes are powers of two 16..256", N);
  if (gen && !gen_size(N))
    return fail(EXMI_ERR_UNSUPPORTED, "%d-D num_points=%d: supported sizes are 2..%d", D, N, GFFT_MAX_N);
  if (C < 1 || d->batch < 1) return fail(EXMI_ERR_INVALID, "num_channels and batch must be >= 1");
  if (d->order < 0 || d->order > 4) return fail(EXMI_ERR_UNSUPPORTED, "Order %d not implemented.", d->order);

  int kind = NL_ZERO;
  bool conv_general = false;
  if (d->order > 0) {
    if (d->nonlin_kind == EXMI_NL_ZERO) {
      kind = NL_POLY;  // the empty polynomial
    } else if (d->nonlin_kind == EXMI_NL_CONVECTION && d->conv_single_channel && D > 1) {
      if (C != 1) return fail(EXMI_ERR_UNSUPPORTED, "single-channel convection needs num_channels == 1");
      kind = NL_GENERAL;  // as in 2-D: the general form's fields, products and scales
      conv_general = true;
    } else if (d->nonlin_kind == EXMI_NL_PROJECTED_CONVECTION ||
               d->nonlin_kind == EXMI_NL_PROJECTED_CONVECTION_KOLMOGOROV) {
      if (D != 3) return fail(EXMI_ERR_INVALID, "the projected convection is three-dimensional");
      if (C != 3) return fail(EXMI_ERR_INVALID, "the projected convection needs num_channels == 3");
      kind = NL_PROJ3D;
    } else if (d->nonlin_kind == EXMI_NL_VORTICITY || d->nonlin_kind == EXMI_NL_KOLMOGOROV) {
      if (D != 2) return fail(EXMI_ERR_INVALID, "Expected num_spatial_dims = 2, got %d.", D);
      if (C != 1) return fail(EXMI_ERR_INVALID, "vorticity plans need num_channels == 1");
      kind = NL_VORT;
    } else {
      int rc = kernel_kind(d, &kind);
      if (rc) return rc;
    }
  }
  if (d->flags & EXMI_FLAG_FOURIER_RESIDENT)
    return fail(EXMI_ERR_UNSUPPORTED, "Fourier-resident rollouts are built for 1-D plans of the fused sizes");
  if (d->lin_batch != 1 && d->lin_batch != d->batch)
    return fail(EXMI_ERR_INVALID, "lin_batch must be 1 or batch (%d), got %d", d->batch, d->lin_batch);
  if (d->lin_channels != 1 && d->lin_channels != C)
    return fail(EXMI_ERR_INVALID, "lin_channels must be 1 or num_channels (%d), got %d", C, d->lin_channels);
  if (!d->host_linear_operator) return fail(EXMI_ERR_INVALID, "host_linear_operator is null");
  if (d->order > 0 && d->num_circle_points < 1) return fail(EXMI_ERR_INVALID, "num_circle_points must be >= 1");
  const bool cplx = d->host_linear_operator_imag != nullptr;
  exmi_plan* p = new exmi_plan();
  p->d = *d;
  p->d.host_linear_operator = nullptr;
  p->d.host_linear_operator_imag = nullptr;
  p->ndim = D;
  p->md = 1;
  p->gen = gen ? 1 : 0;
  p->gnx = D == 3 ? N : 1;
  p->gny = D >= 2 ? N : 1;
  p->M = N / 2;
  p->n3 = (long long)p->gnx * p->gny * N;
  p->nr = (int)((long long)p->gnx * p->gny * (N / 2 + 1));
  p->ncoef = num_coefs_c(d->order, cplx);
  p->cplx = cplx;
  p->ninst = d->batch * C;
  p->kind3 = kind;
  p->kmax = dealias_kmax(N, d->dealiasing_fraction);
  if (d->order > 0 && d->nonlin_kind == EXMI_NL_ZERO) p->d.num_poly = 0;
  if (d->nonlin_kind == EXMI_NL_CAHN_HILLIARD) cahn_hilliard_poly(&p->d);
  if (conv_general) {
    p->prod_nc = d->conv_conservative ? 0 : 1;
    p->d.general_scales[0] = p->prod_nc ? -d->scale : 0.0;
    p->d.general_scales[1] = p->prod_nc ? 0.0 : -d->scale;
    p->d.general_scales[2] = 0.0;
  }
  p->nf3 = kind == NL_GRADNORM ? D
           : (kind == NL_CONV_NONCONS || kind == NL_GENERAL) ? 1 + D
           : kind == NL_PROJ3D ? 2
           : kind == NL_VORT ? 4
                             : 1;
  if (d->nonlin_kind == EXMI_NL_PROJECTED_CONVECTION_KOLMOGOROV) {
    // _projected_convection.py:195-216: injection_scale times the
    // coef_extraction scaling (_spectral.py:403-527) at (0, k_f, 0): N on
    // the zero-wavenumber axes, N/2 on the k_f axis unless k_f is 0 or N/2
    const int kf = d->injection_mode;
    if (kf >= 0 && kf < N / 2) {
      const double sy = kf == 0 ? N : N / 2.0;
      p->inj_mode = kf;
      p->inj_value = (float)(d->injection_scale * (double)N * sy * (double)N);
    }
  }
  if (d->nonlin_kind == EXMI_NL_KOLMOGOROV) {
    // _vorticity_convection.py:166-176: -k_f gamma scaling(coef_extraction)
    // at wavenumber (0, k_f); scaling = N at k = 0 (and at the Nyquist mode
    // of an even N), else N/2 per axis
    const int kf = d->injection_mode;
    if (kf >= 0 && kf <= N / 2) {
      const double sy = (kf == 0 || 2 * kf == N) ? N : N / 2.0;
      p->inj_mode = kf;
      p->inj_value = (float)(-(double)kf * d->injection_scale * (double)N * sy);
    }
  }
  p->np3 = kind == NL_CONV_CONS ? D : kind == NL_GENERAL ? 2 : 1;
  p->kscale = (float)(2.0 * 3.14159265358979323846 / d->domain_extent);
  // Spectrum-carrying rollouts (strategy 2): with a real L and, for an even
  // N, the Nyquist modes dealiased out of the nonlinear term, the forward
  // transform of an emitted state returns the spectrum it came from (exact
  // arithmetic), so the rollout keeps the spectrum from step to step and
  // only inverse-transforms each state it emits; the autotuner times it
  // against the per-step round trip (-1).
  p->md_carry = !cplx && (N % 2 == 1 || p->kmax < N / 2) && !(d->flags & EXMI_FLAG_PHYSICAL_ROLLOUT);
  p->roll_nt = p->md_carry ? 0 : -1;
  hipGetDevice(&p->device);
  int rc = gen ? EXMI_OK : get_twiddles(N / 2, &p->tw);
  if (!rc) rc = get_twiddles(N, &p->tw_col);
  if (rc) {
    delete p;
    return rc;
  }
  if (gen) gfft_cfg_from_code(N, gfft_default_code(), &p->gcfg);
  p->gcode = gen ? gfft_default_code() : 0;
  const long long rows = (long long)d->lin_batch * d->lin_channels, total = rows * p->nr;
  double* dL = nullptr;
  const long long I = p->ninst, n3 = p->n3, nfp = p->nf3 > p->np3 ? p->nf3 : p->np3;
  const bool nl = d->order > 0;
  bool ok = hipMalloc(&dL, sizeof(double) * total * (cplx ? 2 : 1)) == hipSuccess &&
            hipMalloc(&p->tables, sizeof(float) * total * p->ncoef) == hipSuccess &&
            hipMalloc(&p->W3, sizeof(cf) * I * p->nr) == hipSuccess &&
            hipMalloc(&p->P3, sizeof(float) * I * n3) == hipSuccess &&
            (!p->md_carry || hipMalloc(&p->K3, sizeof(cf) * I * p->nr) == hipSuccess) &&
            (!nl || (hipMalloc(&p->N03, sizeof(cf) * I * p->nr) == hipSuccess &&
                     hipMalloc(&p->NC3, sizeof(cf) * I * p->nr) == hipSuccess &&
                     hipMalloc(&p->S3, sizeof(cf) * I * p->nr) == hipSuccess &&
                     hipMalloc(&p->F3, sizeof(cf) * I * nfp * p->nr) == hipSuccess &&
                     hipMalloc(&p->R3, sizeof(float) * I * p->nf3 * n3) == hipSuccess &&
                     hipMalloc(&p->PR3, sizeof(float) * I * p->np3 * (2 * p->nr > n3 ? 2 * p->nr : n3)) ==
                         hipSuccess)) &&
            (d->order < 3 || hipMalloc(&p->A3, sizeof(cf) * I * p->nr) == hipSuccess) &&
            (d->order < 4 || hipMalloc(&p->S13, sizeof(cf) * I * p->nr) == hipSuccess);
  if (!ok) {
    hipFree(dL);
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_ALLOC, "%d-D workspace allocation failed (batch %d, N %d)", D, d->batch, N);
  }
  double* dF = nullptr;
  int frc = upload_nonlin_factor(d, rows, &dF);
  if (frc) {
    hipFree(dL);
    exmi_plan_destroy(p);
    return frc;
  }
  hipError_t e = hipMemcpy(dL, d->host_linear_operator, sizeof(double) * total, hipMemcpyHostToDevice);
  if (e == hipSuccess && cplx)
    e = hipMemcpy(dL + total, d->host_linear_operator_imag, sizeof(double) * total, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_coef(dL, cplx ? dL + total : nullptr, total, (int)p->nr, d->dt, d->order, d->num_circle_points,
                    d->circle_radius, dF, p->tables, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  hipFree(dL);
  hipFree(dF);
  if (e != hipSuccess) {
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_HIP, "coefficient formation failed: %s", hipGetErrorString(e));
  }
  if (hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&p->launch_stream, hipStreamNonBlocking) != hipSuccess) {
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_HIP, "stream creation failed");
  }
  if (gen) {
    rc = tune_gfft(p);
    if (rc) {
      exmi_plan_destroy(p);
      return rc;
    }
  }
  *out = p;
  return EXMI_OK;
}

static hipError_t rfft_md(const exmi_plan* p, long long I, const float* x, cf* X, hipStream_t st) {
  if (p->gen) return rfft_gen(p->gcfg, p->ndim, I, x, X, p->tw_col.tw, st);
  return rfft3(p->d.num_points, I, x, X, p->tw, p->tw_col, st);
}
static hipError_t irfft_md(const exmi_plan* p, long long I, cf* X, float* x, hipStream_t st) {
  if (p->gen) return irfft_gen(p->gcfg, p->ndim, I, X, x, p->tw_col.tw, st);
  return irfft3(p->d.num_points, I, X, x, p->tw, p->tw_col, st);
}
// irfftn of the spectra X that leaves X unchanged (the intermediate in K)
static hipError_t irfft_md_keep(const exmi_plan* p, long long I, const cf* X, cf* K, float* x, hipStream_t st) {
  if (p->gen) return irfft_gen(p->gcfg, p->ndim, I, K, x, p->tw_col.tw, st, X);
  return irfft3(p->d.num_points, I, K, x, p->tw, p->tw_col, st, X);
}
// The same for a fused 3-D size, each state (C N^3 floats) written to x +
// b * xstride (a strided batch such as a trajectory buffer).
static hipError_t irfft3_keep_strided(const exmi_plan* p, const cf* X, cf* K, float* x, long long xstride,
                                      hipStream_t st) {
  const int N = p->d.num_points;
  return irfft3(N, p->ninst, K, x, p->tw, p->tw_col, st, X, (long long)p->d.num_channels * N * N, xstride);
}

static void fill_nl3d(const exmi_plan* p, NL3DArgs& a) {
  const int N = p->d.num_points;
  memset(&a, 0, sizeof a);
  a.N = N;
  a.NZ = N / 2 + 1;
  a.dims = p->ndim;
  a.nx = p->gnx;
  a.ny = p->gny;
  a.channels = p->d.num_channels;
  a.kind = p->kind3;
  a.kmax = p->kmax;
  a.nf = p->nf3;
  a.np = p->np3;
  a.Nr = p->nr;
  a.N3 = p->n3;
  a.ninst = p->ninst;
  a.kscale = p->kscale;
  a.scale = (float)p->d.scale;
  a.g0 = (float)p->d.general_scales[0];
  a.prod_nc = p->prod_nc;
  a.g1 = (float)p->d.general_scales[1];
  a.g2 = (float)p->d.general_scales[2];
  a.feed = (float)p->d.reaction_rates[0];
  a.kill = (float)p->d.reaction_rates[1];
  a.zero_mode_fix = p->d.zero_mode_fix;
  a.npoly = p->d.num_poly;
  for (int i = 0; i < EXMI_MAX_POLY; ++i) a.poly[i] = (float)p->d.poly[i];
  a.poly_lap = p->d.nonlin_kind == EXMI_NL_CAHN_HILLIARD;
  a.inj_mode = p->inj_mode;
  a.inj_axis = p->kind3 == NL_VORT ? 2 : 1;
  a.inj_value = p->inj_value;
  a.nh = p->adj_nh;
  a.nq = p->adj_nq;
}

// Fused 3-D plans whose z rows run with the products fused (zprod3d_kernel).
static bool zprod_kind(const exmi_plan* p) {
  const int k3 = p->kind3;
  return !p->gen && p->d.num_points <= 128 &&
         (k3 == NL_GRADNORM || k3 == NL_GENERAL || k3 == NL_POLY || k3 == NL_GRAY_SCOTT || k3 == NL_BZ ||
          k3 == NL_CONV_CONS || k3 == NL_CONV_NONCONS);
}

// Fused 3-D plans whose forward x lines run with the combine and the stage
// (xstage3d_kernel, xstage_proj3d_kernel) inside an ETDRK step.
static bool xstage_kind(const exmi_plan* p) {
  return zprod_kind(p) || (!p->gen && p->kind3 == NL_PROJ3D && p->d.num_points <= 128);
}

// The product spectra of N(X) of every instance into F3 (the passes of
// spectral3d.hip up to the combine).
// *ph: where the product spectra are (F3, or for the projected convection
// of a fused size R3 as spectra). no_x: the forward x lines of the kinds
// with fused z rows are left to the caller (stages3d fuses them with the
// stage; see xstage_kind).
static hipError_t nonlinear3d_products(exmi_plan* p, const NL3DArgs& a, const cf* X, hipStream_t st,
                                       const cf** ph, bool no_x = false) {
  *ph = p->F3;
  if (p->gen) {
    hipError_t e = launch_fields3d(a, X, p->F3, st);
    if (e == hipSuccess) e = irfft_md(p, p->ninst * p->nf3, p->F3, p->R3, st);
    if (e == hipSuccess) e = launch_products3d(a, p->R3, p->PR3, st);
    if (e == hipSuccess) e = rfft_md(p, p->ninst * p->np3, p->PR3, p->F3, st);
    return e;
  }
  // 3-D grids of the fused sizes. The fields are dealiased and the combine
  // keeps only |k| <= kmax, so every pass skips the kz planes above the
  // cutoff: the inverse x lines (fused with the fields pass) and y lines run
  // on planes kz <= kmax only, the z rows read the planes above as zero; the
  // forward z rows store, and the y and x lines transform, planes kz <= kmax
  // only (the combine never reads the others).
  const int N = p->d.num_points;
  const long long NZ = N / 2 + 1, I = p->ninst * p->nf3, IP = p->ninst * p->np3;
  const int kz = p->kmax < (int)NZ - 1 ? (p->kmax < 0 ? 0 : p->kmax) : (int)NZ - 1;  // last kept plane
  const float r = 1.f / (float)N;
  const cf* tl = p->tw_col.tw;
  // the fields with the inverse x lines: one lane group per trajectory or
  // instance line where the fields share their spectra (each spectrum value
  // read once), else per field line
  const int k3 = p->kind3;
  const bool unit = N <= 128 && (k3 == NL_PROJ3D || k3 == NL_GRADNORM || k3 == NL_GENERAL || k3 == NL_CONV_NONCONS);
  hipError_t e = unit ? launch_fields_unit_lines3d(a, X, p->F3, r, tl, kz + 1, st)
                      : launch_fields_lines3d(a, X, p->F3, r, tl, kz + 1, st);
  // up to N = 64 (2/3 cutoff) the inverse y lines, the z rows with the
  // products and the forward y lines are one plane kernel (yzprod3d_kernel)
  // for the kinds it covers
  const bool yz = N <= 64 && kz < N / 3 + 1 &&
                  (k3 == NL_PROJ3D || k3 == NL_GRADNORM || k3 == NL_GENERAL || k3 == NL_POLY ||
                   k3 == NL_GRAY_SCOTT || k3 == NL_BZ);
  if (e == hipSuccess && !yz)
    e = launch_lines(N, true, p->F3, I * N * NZ, NZ, N * NZ, NZ, r, tl, st, nullptr, kz + 1);
  if (p->kind3 == NL_PROJ3D) {
    // the z rows with the cross products fused (zcross3d_kernel, or with
    // the y lines yzprod3d_kernel), the product spectra in R3 (which holds
    // I N^3 >= ninst N^2 (N + 2) floats)
    cf* P = (cf*)p->R3;
    if (e == hipSuccess)
      e = yz ? launch_yzprod3d(a, p->F3, P, kz, tl, p->tw.rt(), st)
             : launch_zcross3d(N, p->F3, P, p->d.batch, p->nr, kz, p->tw.rt(), st);
    if (e == hipSuccess && !yz)
      e = launch_lines(N, false, P, IP * N * NZ, NZ, N * NZ, NZ, 1.f, tl, st, nullptr, kz + 1);
    if (e == hipSuccess && !no_x)
      e = launch_lines(N, false, P, IP * N * NZ, N * NZ, (long long)N * N * NZ, N * NZ, 1.f, tl, st, nullptr, kz + 1);
    *ph = P;
    return e;
  }
  if (zprod_kind(p)) {
    // the z rows with the products fused (zprod3d_kernel, or with the y
    // lines yzprod3d_kernel); one field and one product per instance: in
    // place in F3; three products per instance
    // (the conservative convection): into PR3 (allocated to hold them as
    // spectra); else into R3 (which holds I NF N^3 >= ninst NP N^2 (N + 2)
    // floats)
    cf* P = p->nf3 == p->np3 ? p->F3 : p->kind3 == NL_CONV_CONS ? (cf*)p->PR3 : (cf*)p->R3;
    if (e == hipSuccess) e = yz ? launch_yzprod3d(a, p->F3, P, kz, tl, p->tw.rt(), st)
                                : launch_zprod3d(a, p->F3, P, kz, p->tw.rt(), st);
    if (e == hipSuccess && !yz)
      e = launch_lines(N, false, P, IP * N * NZ, NZ, N * NZ, NZ, 1.f, tl, st, nullptr, kz + 1);
    if (e == hipSuccess && !no_x)
      e = launch_lines(N, false, P, IP * N * NZ, N * NZ, (long long)N * N * NZ, N * NZ, 1.f, tl, st, nullptr, kz + 1);
    *ph = P;
    return e;
  }
  if (e == hipSuccess)
    e = launch_fft1d(N / 2, true, (const float*)p->F3, p->R3, (int)(I * N * N), p->tw.rt(), st, p->kmax);
  if (e == hipSuccess) e = launch_products3d(a, p->R3, p->PR3, st);
  if (e == hipSuccess) e = launch_fft1d(N / 2, false, p->PR3, (float*)p->F3, (int)(IP * N * N), p->tw.rt(), st, p->kmax);
  if (e == hipSuccess) e = launch_lines(N, false, p->F3, IP * N * NZ, NZ, N * NZ, NZ, 1.f, tl, st, nullptr, kz + 1);
  if (e == hipSuccess)
    e = launch_lines(N, false, p->F3, IP * N * NZ, N * NZ, (long long)N * N * NZ, N * NZ, 1.f, tl, st, nullptr, kz + 1);
  return e;
}

// N(X) of every instance into outN.
static hipError_t nonlinear3d(exmi_plan* p, const cf* X, cf* outN, hipStream_t st) {
  NL3DArgs a;
  fill_nl3d(p, a);
  const cf* ph = nullptr;
  hipError_t e = nonlinear3d_products(p, a, X, st, &ph);
  if (e == hipSuccess) e = launch_combine3d(a, ph, outN, st);
  return e;
}

// The stages of one ETDRK1-4 step on the spectra in sa (U0 the state, S the
// result): each stage follows the evaluation of N at its argument (U0 for
// the first, then the previous stage's S), the combine pass fused into the
// stage (the first also keeps N0).
// For the fused sizes up to 128 the forward x lines of the product spectra
// run fused with the combine (the projection) and the stage (xstage_kind).
// With emitK (a rollout that emits the new state) the last stage also
// stores the new state's inverse x lines to emitK (the first pass of
// irfftn); *emitted says whether it did.
static hipError_t stages3d(exmi_plan* p, Stage3DArgs& sa, hipStream_t st, cf* emitK = nullptr,
                           bool* emitted = nullptr) {
  static const int seq[4][4] = {{ST1, -1}, {ST2A, ST2B, -1}, {ST3A, ST3B, ST3C, -1}, {ST4A, ST4B, ST4C, ST4D}};
  NL3DArgs a;
  fill_nl3d(p, a);
  const int N = p->d.num_points, NZ = N / 2 + 1;
  const bool xs = xstage_kind(p);
  const int kz = p->kmax < NZ - 1 ? (p->kmax < 0 ? 0 : p->kmax) : NZ - 1;
  const bool em = emitK && xs;
  if (emitted) *emitted = em;
  hipError_t e = hipSuccess;
  for (int i = 0; i < 4 && e == hipSuccess; ++i) {
    const int m = seq[p->d.order - 1][i];
    if (m < 0) break;
    const bool last = i == 3 || seq[p->d.order - 1][i + 1] < 0;
    const cf* ph = nullptr;
    e = nonlinear3d_products(p, a, i == 0 ? sa.U0 : sa.S, st, &ph, xs);
    sa.mode = m;
    cf* n0 = i == 0 ? p->N03 : nullptr;
    if (e == hipSuccess)
      e = !xs                          ? launch_stage_nl3d(a, sa, ph, n0, st)
          : p->kind3 == NL_PROJ3D ? launch_xstage_proj3d(a, sa, ph, n0, kz, p->tw_col.tw, st,
                                                         em && last ? emitK : nullptr)
                                  : launch_xstage3d(a, sa, ph, n0, kz, p->tw_col.tw, st, em && last ? emitK : nullptr);
  }
  return e;
}

// One multi-pass step u_in (batch stride ibs) -> u_out (obs): forward
// transform, ETDRK stages, inverse transform; strided states pass through
// the staging buffer.
// With in_f / out_f the input / output are spectra (C, N^(D-1), N/2+1) with
// batch strides in complex entries (exmi_step_fourier); nonlin_only
// evaluates N(u_hat) instead of a step (exmi_nonlinear).
static int enqueue_step3d(exmi_plan* p, const float* u_in, long long ibs, float* u_out, long long obs,
                          hipStream_t st, bool in_f = false, bool out_f = false, bool nonlin_only = false) {
  const int B = p->d.batch;
  const long long sz = (long long)p->d.num_channels * p->n3;
  const long long szs = 2LL * p->d.num_channels * p->nr;  // floats of one spectrum
  hipError_t e = hipSuccess;
  if (in_f) {
    e = copy_batch(u_in, 2 * ibs, (float*)p->W3, szs, szs, B, st);
  } else {
    const float* src = u_in;
    if (ibs != sz) {
      e = copy_batch(u_in, ibs, p->P3, sz, sz, B, st);
      src = p->P3;
    }
    if (e == hipSuccess) e = rfft_md(p, p->ninst, src, p->W3, st);
  }
  cf* X = p->W3;  // the new state's spectrum
  if (nonlin_only) {
    if (e == hipSuccess) e = nonlinear3d(p, p->W3, p->N03, st);
    X = p->N03;
  } else if (p->d.order == 0) {
    if (e == hipSuccess)
      e = launch_etdrk0_spec(p->W3, p->ninst, p->nr, p->d.num_channels, p->tables, p->ncoef, p->d.lin_batch,
                             p->d.lin_channels, p->cplx ? num_coefs(0) : -1, st);
  } else {
    Stage3DArgs sa;
    sa.ninst = p->ninst;
    sa.nr = p->nr;
    sa.channels = p->d.num_channels;
    sa.ncoef = p->ncoef;
    sa.lin_batch = p->d.lin_batch;
    sa.lin_channels = p->d.lin_channels;
    sa.e_im = p->cplx ? num_coefs(p->d.order) : -1;
    sa.tables = p->tables;
    sa.U0 = p->W3;
    sa.N0 = p->N03;
    sa.NC = p->NC3;
    sa.S = p->S3;
    sa.S1 = p->S13;
    sa.A = p->A3;
    // etdrk/_etdrk_{1,2,3,4}.py: each stage after the evaluation of N at its argument
    if (e == hipSuccess) e = stages3d(p, sa, st);
    X = p->S3;
  }
  if (out_f) {
    if (e == hipSuccess) e = copy_batch((const float*)X, szs, u_out, 2 * obs, szs, B, st);
  } else {
    float* dst = obs == sz ? u_out : p->P3;
    if (e == hipSuccess) e = irfft_md(p, p->ninst, X, dst, st);
    if (e == hipSuccess && dst != u_out) e = copy_batch(p->P3, sz, u_out, obs, sz, B, st);
  }
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "%d-D step launch failed: %s", p->ndim, hipGetErrorString(e));
}

// Multi-pass spectrum-carrying rollout (strategy 2): the initial states are
// transformed once; each step advances the spectra (the stages of
// enqueue_step3d, alternating the state between W3 and S3) and emits its
// state with the inverse transform that leaves the spectrum in place. States
// go to out + i * out_bs + t * out_ts (only the last, at out, with
// final_only).
static int enqueue_md_carry(exmi_plan* p, const float* u0, long long u0_bs, float* out, long long out_bs,
                            long long out_ts, int n, int final_only, hipStream_t st) {
  const int B = p->d.batch;
  const long long sz = (long long)p->d.num_channels * p->n3;
  hipError_t e = hipSuccess;
  const float* src = u0;
  if (u0_bs != sz) {
    e = copy_batch(u0, u0_bs, p->P3, sz, sz, B, st);
    src = p->P3;
  }
  if (e == hipSuccess) e = rfft_md(p, p->ninst, src, p->W3, st);
  cf* cur = p->W3;
  cf* nxt = p->S3;
  Stage3DArgs sa;
  sa.ninst = p->ninst;
  sa.nr = p->nr;
  sa.channels = p->d.num_channels;
  sa.ncoef = p->ncoef;
  sa.lin_batch = p->d.lin_batch;
  sa.lin_channels = p->d.lin_channels;
  sa.e_im = -1;  // real L
  sa.tables = p->tables;
  sa.N0 = p->N03;
  sa.NC = p->NC3;
  sa.S1 = p->S13;
  sa.A = p->A3;
  for (int t = 0; t < n && e == hipSuccess; ++t) {
    if (p->d.order == 0) {
      e = launch_etdrk0_spec(cur, p->ninst, p->nr, p->d.num_channels, p->tables, p->ncoef, p->d.lin_batch,
                             p->d.lin_channels, -1, st);
    }
    const bool emit = !(final_only && t < n - 1);
    bool emitted = false;  // the last stage left the new state's x lines in K3
    if (p->d.order > 0) {
      sa.U0 = cur;
      sa.S = nxt;
      // up to N = 64 the emitted state's remaining passes are one plane kernel
      e = stages3d(p, sa, st, emit && !p->gen && p->d.num_points <= 64 ? p->K3 : nullptr, &emitted);
      cf* t2 = cur;
      cur = nxt;
      nxt = t2;
    }
    if (e != hipSuccess || !emit) continue;
    float* dst = final_only ? out : out + (long long)t * out_ts;
    if (emitted) {
      const int N = p->d.num_points;
      e = launch_yz_inverse3d(N, p->K3, dst, p->ninst * N, p->tw_col.tw, p->tw.rt(),
                              (long long)p->d.num_channels * N * N, out_bs, st);
      continue;
    }
    if (!p->gen) {
      if (e == hipSuccess) e = irfft3_keep_strided(p, cur, p->K3, dst, out_bs, st);
      continue;
    }
    const bool direct = out_bs == sz;
    if (e == hipSuccess) e = irfft_md_keep(p, p->ninst, cur, p->K3, direct ? dst : p->P3, st);
    if (e == hipSuccess && !direct) e = copy_batch(p->P3, sz, dst, out_bs, sz, B, st);
  }
  return e == hipSuccess ? EXMI_OK
                         : fail(EXMI_ERR_HIP, "%d-D rollout launch failed: %s", p->ndim, hipGetErrorString(e));
}

// ------------------------------------------------------- multi-pass adjoint
// The transposed ETDRK0-4 step on whole spectra (etdrk/_etdrk_{0,1,2,3,4}.py;
// the reverse-mode derivative jax.grad takes of step_fourier), for every
// multi-pass plan whose nonlinear kind has a transpose (adj3d_fields), and,
// through a multi-pass twin of the same stepper, for fused plans the fused
// adjoint kernels do not cover (1-D ETDRK3/4, 2-D kinds other than the
// gradient norm, 2-D ETDRK orders other than 2). With lam the cotangent of
// the new state and J(x)^T the nonlinear transpose at x (all multipliers
// conjugated; the phi-coefficients c_i are real):
//   ETDRK1: E lam + J(u)^T (c1 lam)
//   ETDRK2 (ua = E u + c1 N(u)): rho = lam + J(ua)^T (c2 lam),
//           E rho + J(u)^T (c1 rho - c2 lam)
//   ETDRK3 (ua = Eh u + c2 N0, ub = E u + c3 (2 N(ua) - N0),
//           u' = E u + c4 N0 + c5 N(ua) + c6 N(ub); table indices):
//           l_ub = J(ub)^T (c6 lam), l_ua = J(ua)^T (c5 lam + 2 c3 l_ub),
//           E (lam + l_ub) + Eh l_ua + J(u)^T (c4 lam - c3 l_ub + c2 l_ua)
//   ETDRK4 (ua = Eh u + c2 N0, ub = Eh u + c2 N(ua), uc = Eh ua + c2 (2 N(ub) - N0),
//           u' = E u + c3 N0 + 2 c4 (N(ua) + N(ub)) + c5 N(uc)):
//           l_uc = J(uc)^T (c5 lam), l_ub = J(ub)^T (2 c4 lam + 2 c2 l_uc),
//           l_ua = Eh l_uc + J(ua)^T (2 c4 lam + c2 l_ub),
//           E lam + Eh (l_ub + l_ua) + J(u)^T (c3 lam - c2 l_uc + c2 l_ua)
// The stage states ua, ub, uc are recomputed from u with the forward stage
// kernels.
static bool md_adjoint_ok(const exmi_plan* p) {
  if (p->d.order == 0) return true;
  int nh = 0, nq = 0;
  return adj3d_fields(p->kind3, p->ndim, p->d.nonlin_kind == EXMI_NL_CAHN_HILLIARD, &nh, &nq) > 0;
}

static int ensure_adj_md(exmi_plan* p) {
  if (p->LAM3) return EXMI_OK;
  int nh = 0, nq = 0;
  const int order = p->d.order;
  if (order > 0) adj3d_fields(p->kind3, p->ndim, p->d.nonlin_kind == EXMI_NL_CAHN_HILLIARD, &nh, &nq);
  p->adj_nh = nh;
  p->adj_nq = nq;
  const long long I = p->ninst, sp = I * p->nr, nfh = nh > nq ? nh : nq;
  auto al = [&](cf** b, bool need) { return !need || hipMalloc(b, sizeof(cf) * sp) == hipSuccess; };
  bool ok = al(&p->LAM3, true) && al(&p->MU3, true) && al(&p->T3, true) && al(&p->XA3, order >= 2) &&
            al(&p->LA3, order >= 2) && al(&p->XB3, order >= 3) && al(&p->LB3, order >= 3) &&
            al(&p->XC3, order >= 4) && al(&p->LC3, order >= 4) &&
            (order == 0 || (hipMalloc(&p->FH3, sizeof(cf) * nfh * sp) == hipSuccess &&
                            hipMalloc(&p->RH3, sizeof(float) * I * nh * p->n3) == hipSuccess &&
                            hipMalloc(&p->Q3, sizeof(float) * I * nq * p->n3) == hipSuccess));
  if (!ok) {
    for (cf** b : {&p->XA3, &p->XB3, &p->XC3, &p->LAM3, &p->MU3, &p->T3, &p->LA3, &p->LB3, &p->LC3, &p->FH3}) {
      hipFree(*b);
      *b = nullptr;
    }
    hipFree(p->RH3);
    hipFree(p->Q3);
    p->RH3 = p->Q3 = nullptr;
    return fail(EXMI_ERR_ALLOC, "adjoint workspace allocation failed (batch %d, N %d)", p->d.batch,
                p->d.num_points);
  }
  return EXMI_OK;
}

// out = J(W)^T MU for every instance (out aliases neither W nor MU).
static hipError_t nonlinear_vjp3d(exmi_plan* p, const cf* W, const cf* MU, cf* out, hipStream_t st) {
  NL3DArgs a;
  fill_nl3d(p, a);
  hipError_t e = launch_fields3d(a, W, p->F3, st);
  if (e == hipSuccess) e = irfft_md(p, p->ninst * p->nf3, p->F3, p->R3, st);
  if (e == hipSuccess) e = launch_adj_cofields3d(a, MU, p->FH3, st);
  if (e == hipSuccess) e = irfft_md(p, p->ninst * p->adj_nh, p->FH3, p->RH3, st);
  if (e == hipSuccess) e = launch_adj_products3d(a, p->R3, p->RH3, p->Q3, st);
  if (e == hipSuccess) e = rfft_md(p, p->ninst * p->adj_nq, p->Q3, p->FH3, st);
  if (e == hipSuccess) e = launch_adj_combine3d(a, p->FH3, out, st);
  return e;
}

struct AdjTerm {
  const cf* x;
  int op;  // -1 identity, 100 conj(E), 101 conj(E_half), i >= 0 the real coefficient i
  float mult;
};

static hipError_t adj_lin(const exmi_plan* p, cf* out, std::initializer_list<AdjTerm> terms, hipStream_t st) {
  AdjLinArgs a;
  memset(&a, 0, sizeof a);
  a.ninst = p->ninst;
  a.nr = p->nr;
  a.channels = p->d.num_channels;
  a.ncoef = p->ncoef;
  a.lin_batch = p->d.lin_batch;
  a.lin_channels = p->d.lin_channels;
  a.e_im = p->cplx ? num_coefs(p->d.order) : -1;
  a.tables = p->tables;
  if (terms.size() > 4) return hipErrorInvalidValue;
  for (const AdjTerm& t : terms) {
    const bool ok = t.op == -1 || (t.op >= 0 && t.op < num_coefs(p->d.order)) || t.op == 100 ||
                    (t.op == 101 && p->d.order >= 3);
    if (!ok) return hipErrorInvalidValue;  // would read outside the tables
    a.x[a.nt] = t.x;
    a.op[a.nt] = t.op;
    a.mult[a.nt] = t.mult;
    ++a.nt;
  }
  a.out = out;
  return launch_adj_linear3d(a, st);
}

static int enqueue_vjp_md(exmi_plan* p, const float* u, long long ubs, const float* g, long long gbs, float* gout,
                          long long obs, hipStream_t st) {
  const int B = p->d.batch, order = p->d.order;
  const long long sz = (long long)p->d.num_channels * p->n3;
  hipError_t e = hipSuccess;
  auto to_spec = [&](const float* x, long long xbs, cf* X) {
    const float* src = x;
    if (e == hipSuccess && xbs != sz) {
      e = copy_batch(x, xbs, p->P3, sz, sz, B, st);
      src = p->P3;
    }
    if (e == hipSuccess) e = rfft_md(p, p->ninst, src, X, st);
  };
  to_spec(u, ubs, p->W3);
  to_spec(g, gbs, p->LAM3);
  cf *U = p->W3, *LAM = p->LAM3, *MU = p->MU3, *T = p->T3;
  auto vjp = [&](const cf* W, const cf* M, cf* out) {
    if (e == hipSuccess) e = nonlinear_vjp3d(p, W, M, out, st);
  };
  auto lin = [&](cf* out, std::initializer_list<AdjTerm> terms) {
    if (e == hipSuccess) e = adj_lin(p, out, terms, st);
  };
  if (order >= 2) {  // the forward stage states
    Stage3DArgs sa;
    sa.ninst = p->ninst;
    sa.nr = p->nr;
    sa.channels = p->d.num_channels;
    sa.ncoef = p->ncoef;
    sa.lin_batch = p->d.lin_batch;
    sa.lin_channels = p->d.lin_channels;
    sa.e_im = p->cplx ? num_coefs(order) : -1;
    sa.tables = p->tables;
    sa.U0 = U;
    sa.N0 = p->N03;
    sa.NC = p->NC3;
    sa.S1 = p->S13;
    sa.A = p->A3;
    auto stage = [&](int mode, cf* S) {
      sa.S = S;
      sa.mode = mode;
      if (e == hipSuccess) e = launch_stage3d(sa, st);
    };
    auto nlev = [&](const cf* in, cf* outN) {
      if (e == hipSuccess) e = nonlinear3d(p, in, outN, st);
    };
    nlev(U, p->N03);
    if (order == 2) {
      stage(ST2A, p->XA3);
    } else if (order == 3) {
      stage(ST3A, p->XA3);
      nlev(p->XA3, p->NC3);
      stage(ST3B, p->XB3);
    } else {
      stage(ST4A, p->XA3);
      nlev(p->XA3, p->NC3);
      stage(ST4B, p->XB3);
      nlev(p->XB3, p->NC3);
      stage(ST4C, p->XC3);
    }
  }
  cf* OUT = MU;  // the last combination reads neither MU nor OUT
  switch (order) {
    case 0: lin(OUT, {{LAM, 100, 1.f}}); break;
    case 1:
      lin(MU, {{LAM, 1, 1.f}});
      vjp(U, MU, T);
      lin(OUT, {{LAM, 100, 1.f}, {T, -1, 1.f}});
      break;
    case 2: {
      cf* RHO = p->LA3;
      lin(MU, {{LAM, 2, 1.f}});
      vjp(p->XA3, MU, T);
      lin(RHO, {{LAM, -1, 1.f}, {T, -1, 1.f}});
      lin(MU, {{RHO, 1, 1.f}, {LAM, 2, -1.f}});
      vjp(U, MU, T);
      lin(OUT, {{RHO, 100, 1.f}, {T, -1, 1.f}});
      break;
    }
    case 3:
      lin(MU, {{LAM, 6, 1.f}});
      vjp(p->XB3, MU, p->LB3);
      lin(MU, {{LAM, 5, 1.f}, {p->LB3, 3, 2.f}});
      vjp(p->XA3, MU, p->LA3);
      lin(MU, {{LAM, 4, 1.f}, {p->LB3, 3, -1.f}, {p->LA3, 2, 1.f}});
      vjp(U, MU, T);
      lin(OUT, {{LAM, 100, 1.f}, {p->LB3, 100, 1.f}, {p->LA3, 101, 1.f}, {T, -1, 1.f}});
      break;
    default:
      lin(MU, {{LAM, 5, 1.f}});
      vjp(p->XC3, MU, p->LC3);
      lin(MU, {{LAM, 4, 2.f}, {p->LC3, 2, 2.f}});
      vjp(p->XB3, MU, p->LB3);
      lin(MU, {{LAM, 4, 2.f}, {p->LB3, 2, 1.f}});
      vjp(p->XA3, MU, T);
      lin(p->LA3, {{p->LC3, 101, 1.f}, {T, -1, 1.f}});
      lin(MU, {{LAM, 3, 1.f}, {p->LC3, 2, -1.f}, {p->LA3, 2, 1.f}});
      vjp(U, MU, T);
      lin(OUT, {{LAM, 100, 1.f}, {p->LB3, 101, 1.f}, {p->LA3, 101, 1.f}, {T, -1, 1.f}});
      break;
  }
  float* dst = obs == sz ? gout : p->P3;
  if (e == hipSuccess) e = irfft_md(p, p->ninst, OUT, dst, st);
  if (e == hipSuccess && dst != gout) e = copy_batch(p->P3, sz, gout, obs, sz, B, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "adjoint step launch failed: %s", hipGetErrorString(e));
}

// Fused plans without a fused adjoint build a multi-pass twin of the same
// stepper (same description and operator tables) the first time an adjoint
// is asked for; synchronous.
static bool twin_kind_ok(const exmi_plan_desc* d) {
  if (d->order == 0) return true;
  switch (d->nonlin_kind) {
    case EXMI_NL_ZERO:
    case EXMI_NL_CONVECTION:
    case EXMI_NL_GRADIENT_NORM:
    case EXMI_NL_POLYNOMIAL:
    case EXMI_NL_GENERAL:
    case EXMI_NL_VORTICITY:
    case EXMI_NL_KOLMOGOROV:
    case EXMI_NL_CAHN_HILLIARD:
    case EXMI_NL_GRAY_SCOTT:
    case EXMI_NL_BELOUSOV_ZHABOTINSKY: return true;
    default: return false;
  }
}

// Keep what the twin needs (called at the end of a fused plan's creation).
static void keep_for_twin(exmi_plan* p, const exmi_plan_desc* d, long long total) {
  if (!twin_kind_ok(d)) return;
  p->d0 = *d;
  p->d0.host_linear_operator = nullptr;
  p->d0.host_linear_operator_imag = nullptr;
  p->d0.host_nonlin_factor = nullptr;
  p->hL.assign(d->host_linear_operator, d->host_linear_operator + total);
  if (d->host_linear_operator_imag)
    p->hLi.assign(d->host_linear_operator_imag, d->host_linear_operator_imag + total);
  if (d->host_nonlin_factor) p->hF.assign(d->host_nonlin_factor, d->host_nonlin_factor + d->batch);
}

static int ensure_twin(exmi_plan* p) {
  if (p->twin) return EXMI_OK;
  if (p->hL.empty()) return fail(EXMI_ERR_UNSUPPORTED, "no adjoint for this plan");
  exmi_plan_desc d = p->d0;
  d.host_linear_operator = p->hL.data();
  d.host_linear_operator_imag = p->hLi.empty() ? nullptr : p->hLi.data();
  d.host_nonlin_factor = p->hF.empty() ? nullptr : p->hF.data();
  d.flags &= ~(EXMI_FLAG_FOURIER_RESIDENT | EXMI_FLAG_PHYSICAL_ROLLOUT);
  exmi_plan* t = nullptr;
  int rc = plan_create_md(&d, &t, true);
  if (rc) return rc;
  if (!md_adjoint_ok(t)) {
    exmi_plan_destroy(t);
    return fail(EXMI_ERR_UNSUPPORTED, "no adjoint for this plan's nonlinear function");
  }
  rc = ensure_adj_md(t);
  if (rc) {
    exmi_plan_destroy(t);
    return rc;
  }
  p->twin = t;
  return EXMI_OK;
}

static int enqueue_step2d(exmi_plan* p, const float* u_in, long long ibs, float* u_out, long long obs, bool have_T,
                          bool write_T, hipStream_t st);
static int transform2d(bool inverse, int N, int count, const float* x, cf* xh, hipStream_t st);

// 2-D step_fourier (_base_stepper.py:222-239, etdrk/_base_etdrk.py:64-80):
// the step chain's column passes from the given spectrum to the new one, no
// outer transforms: the first column pass reads u_hat [kx][ky] (COL_CA_SPEC,
// or COL_C0 for ETDRK0) and the last stage writes u_hat' [kx][ky]; the row
// passes in between form the nonlinear products as in a physical step.
// Imaginary parts of the DC and Nyquist columns are carried like every other
// mode, as the reference's step_fourier does. Contiguous spectra; in and out
// may be the same buffer (every read of `in` happens in the first launch,
// every write of `out` in the last).
static int step_fourier2d(exmi_plan* p, const cf* in, long long ibs, cf* out, long long obs, hipStream_t st) {
  const int N = p->d.num_points;
  const long long nr = (long long)N * (N / 2 + 1), szs = (long long)p->d.num_channels * nr;
  if (ibs != szs || obs != szs) return fail(EXMI_ERR_UNSUPPORTED, "exmi_step_fourier: 2-D spectra must be contiguous");
  Step2DArgs a;
  fill_args2d(p, &a);
  a.spec = const_cast<cf*>(in);
  a.spec_out = out;
  a.spec_bstride = nr;
  a.spec_io = 1;
  const int B = p->ninst;
  hipError_t e = hipSuccess;
  if (p->d.order == 0) {
    e = p->col2d[COL_C0](a, B, st);
  } else if (p->d.order == 1) {
    e = p->col2d[COL_CA_SPEC](a, B, st);
    if (e == hipSuccess) e = p->row2d[ROW_RB](a, B, st);
    if (e == hipSuccess) e = p->col2d[COL_C1](a, B, st);
  } else {
    static const int stages[3][4] = {{COL_CB, COL_CC, -1, -1},
                                     {COL_C3A, COL_C3B, COL_C3C, -1},
                                     {COL_C4A, COL_C4B, COL_C4C, COL_C4D}};
    e = p->col2d[COL_CA_SPEC](a, B, st);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) {
      const int m = stages[p->d.order - 2][k];
      if (m < 0) break;
      e = p->row2d[ROW_RB](a, B, st);
      if (e == hipSuccess) e = p->col2d[m](a, B, st);
    }
  }
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D step_fourier launch failed: %s", hipGetErrorString(e));
}

static int enqueue_step2d(exmi_plan* p, const float* u_in, long long ibs, float* u_out, long long obs, bool have_T,
                          bool write_T, hipStream_t st) {
  Step2DArgs a;
  fill_args2d(p, &a);
  a.u_in = u_in;
  a.u_in_bstride = ibs;
  a.u_out = u_out;
  a.u_out_bstride = obs;
  a.write_T = write_T ? 1 : 0;
  const int B = p->ninst;
  hipError_t e = hipSuccess;
  if (!have_T) e = p->row2d[ROW_RA](a, B, st);
  if (p->d.order == 0) {
    if (e == hipSuccess) e = p->col2d[COL_C0](a, B, st);
  } else if (p->d.order == 1) {
    if (e == hipSuccess) e = p->col2d[COL_CA](a, B, st);
    if (e == hipSuccess) e = p->row2d[ROW_RB](a, B, st);
    if (e == hipSuccess) e = p->col2d[COL_C1](a, B, st);
  } else {
    // ETDRK2 CA RB CB RB CC; ETDRK3 CA RB C3A RB C3B RB C3C;
    // ETDRK4 CA RB C4A RB C4B RB C4C RB C4D
    static const int stages[3][4] = {{COL_CB, COL_CC, -1, -1},
                                     {COL_C3A, COL_C3B, COL_C3C, -1},
                                     {COL_C4A, COL_C4B, COL_C4C, COL_C4D}};
    if (e == hipSuccess) e = p->col2d[COL_CA](a, B, st);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) {
      const int m = stages[p->d.order - 2][k];
      if (m < 0) break;
      e = p->row2d[ROW_RB](a, B, st);
      if (e == hipSuccess) e = p->col2d[m](a, B, st);
    }
  }
  if (e == hipSuccess) e = p->row2d[ROW_RC](a, B, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D step launch failed: %s", hipGetErrorString(e));
}

extern "C" {

const char* exmi_last_error(void) { return g_err.c_str(); }

int exmi_info(char* buf, int32_t buflen) {
  int dev = 0, n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    snprintf(buf, buflen, "exmi: no HIP device");
    return EXMI_OK;
  }
  hipGetDevice(&dev);
  hipDeviceProp_t prop;
  hipGetDeviceProperties(&prop, dev);
  snprintf(buf, buflen, "exmi: device %d %s (%s), %d CUs", dev, prop.name, prop.gcnArchName,
           prop.multiProcessorCount);
  return EXMI_OK;
}

int exmi_plan_create(const exmi_plan_desc* d, exmi_plan** out) {
  g_err.clear();
  if (!d || !out) return fail(EXMI_ERR_INVALID, "null argument");
  *out = nullptr;
  const int D = d->num_spatial_dims, N = d->num_points;
  if (d->host_nonlin_factor && d->lin_batch != d->batch)
    return fail(EXMI_ERR_INVALID, "host_nonlin_factor needs per-trajectory tables (lin_batch == batch)");
  if (D == 2) return size_2d(N) ? plan_create_2d(d, out) : plan_create_md(d, out, true);
  if (D == 3) return plan_create_md(d, out, !size_3d(N));
  if (D != 1) return fail(EXMI_ERR_UNSUPPORTED, "num_spatial_dims=%d: this build supports 1, 2 and 3", D);
  if (!size_1d(N)) return plan_create_md(d, out, true);
  if (d->num_channels < 1 || d->batch < 1) return fail(EXMI_ERR_INVALID, "num_channels and batch must be >= 1");
  if (d->order < 0 || d->order > 4) return fail(EXMI_ERR_UNSUPPORTED, "Order %d not implemented.", d->order);
  if (d->lin_batch != 1 && d->lin_batch != d->batch)
    return fail(EXMI_ERR_INVALID, "lin_batch must be 1 or batch (%d), got %d", d->batch, d->lin_batch);
  if (d->lin_channels != 1 && d->lin_channels != d->num_channels)
    return fail(EXMI_ERR_INVALID, "lin_channels must be 1 or num_channels (%d), got %d", d->num_channels,
                d->lin_channels);
  if (!d->host_linear_operator) return fail(EXMI_ERR_INVALID, "host_linear_operator is null");
  if (d->order > 0 && d->num_circle_points < 1) return fail(EXMI_ERR_INVALID, "num_circle_points must be >= 1");
  const bool cplx = d->host_linear_operator_imag != nullptr;
  if (cplx && N > 1024)
    return fail(EXMI_ERR_UNSUPPORTED, "complex linear operators are built for N <= 1024");
  int kind = NL_ZERO;
  if (d->order > 0) {
    int rc = kernel_kind(d, &kind);
    if (rc) return rc;
  }
  const int kmax = dealias_kmax(N, d->dealiasing_fraction);
  int nl_kind = kind;  // kernel used for exmi_nonlinear: always the literal form
  // Alias-free quadratic products (3 kmax < N, C = D = 1): u u_x and
  // (1/2)(u^2)_x agree on every retained mode, see EXMI_FLAG_EXACT_FORM.
  if (kind == NL_CONV_NONCONS && !(d->flags & EXMI_FLAG_EXACT_FORM) && d->num_channels == 1 && kmax >= 0 &&
      3LL * kmax < N)
    kind = NL_CONV_CONS;
  exmi_plan* p = new exmi_plan();
  p->d = *d;
  p->d.host_linear_operator = nullptr;
  p->d.host_linear_operator_imag = nullptr;
  if (d->nonlin_kind == EXMI_NL_CAHN_HILLIARD) cahn_hilliard_poly(&p->d);
  p->M = N / 2;
  p->nr = p->M + 1;
  p->ncoef = num_coefs_c(d->order, cplx);
  p->kind = kind;
  p->kmax = kmax;
  p->kscale = (float)(2.0 * 3.14159265358979323846 / d->domain_extent);
  hipGetDevice(&p->device);
  // Channel-coupled reactions: every step runs through the multi-step kernel
  // with a trajectory's channels in one lane group (no single-step kernel)
  const bool coupled = kind == NL_GRAY_SCOTT || kind == NL_BZ;
  if (coupled && (cplx || p->M > 512 || (d->flags & EXMI_FLAG_FOURIER_RESIDENT))) {
    delete p;
    return fail(EXMI_ERR_UNSUPPORTED,
                "channel-coupled reactions are built for real linear operators, N <= 1024, without the "
                "Fourier-resident flag");
  }
  p->coupled = coupled;
  p->step = coupled ? nullptr : cplx ? get_step1d_c(p->M, kind, d->order) : get_step1d(p->M, kind, d->order);
  p->nonlin = coupled ? nullptr : d->order > 0 ? get_step1d(p->M, nl_kind, 1) : get_step1d(p->M, NL_ZERO, 1);
  if (!coupled && (!p->step || !p->nonlin)) {
    delete p;
    return fail(EXMI_ERR_UNSUPPORTED, "no kernel for N=%d kind=%d order=%d", N, kind, d->order);
  }
  // the fused 1-D transpose has no Laplacian of Cahn-Hilliard's cubic (the
  // multi-pass twin covers it)
  p->vjp = d->order <= 2 && d->nonlin_kind != EXMI_NL_CAHN_HILLIARD
               ? get_vjp1d(p->M, d->order == 0 ? NL_ZERO : kind, d->order, cplx ? 1 : 0)
               : nullptr;
  if ((d->flags & EXMI_FLAG_FOURIER_RESIDENT) && (d->flags & EXMI_FLAG_PHYSICAL_ROLLOUT)) {
    delete p;
    return fail(EXMI_ERR_INVALID, "EXMI_FLAG_FOURIER_RESIDENT and EXMI_FLAG_PHYSICAL_ROLLOUT exclude each other");
  }
  const int fr_forced = (d->flags & EXMI_FLAG_FOURIER_RESIDENT) ? 1 : 0;
  if (fr_forced && d->order != 2) {
    delete p;
    return fail(EXMI_ERR_UNSUPPORTED, "Fourier-resident rollouts are built for ETDRK2");
  }
  // The forward transform of an emitted state returns the spectrum it came
  // from exactly (in exact arithmetic) when every spectrum the step produces
  // has real DC and Nyquist bins: a real L (real E at both), and a nonlinear
  // term whose Nyquist mode the dealias mask removes (the DC bin of every
  // kind is real or zero). Such ETDRK2 rollouts keep the spectrum on chip.
  const bool fr_exact = !cplx && !coupled && d->order == 2 && (kind == NL_ZERO || kmax < p->M);
  const int fr = fr_forced ? 1 : (!(d->flags & EXMI_FLAG_PHYSICAL_ROLLOUT) && fr_exact) ? 1 : 0;
  p->fourier_resident = fr;
  p->fr_auto = fr && !fr_forced;
  const int std_kc = p->kmax == default_kmax(p->M) ? 1 : 0;
  p->roll[1] = coupled ? get_roll1d_r(p->M, kind, d->order)
              : cplx  ? get_roll1d_c(p->M, kind, d->order, fr)
                      : get_roll1d(p->M, kind, d->order, 1, fr, std_kc);
  p->roll[2] = (cplx || coupled) ? nullptr : get_roll1d(p->M, kind, d->order, 2, fr, std_kc);
  if (const char* e = getenv("EXMI_ROLL_NT")) p->roll_nt = atoi(e);  // 1, 2; -1 = per-step launches
  if (p->fourier_resident && !p->fr_auto && p->roll_nt < 0) p->roll_nt = 0;
  if (coupled) {
    if (!p->roll[1]) {
      delete p;
      return fail(EXMI_ERR_UNSUPPORTED, "no coupled-reaction kernel for N=%d order=%d", N, d->order);
    }
    p->roll_nt = 1;  // the only strategy
  }

  int rc = get_twiddles(p->M, &p->tw);
  if (rc) {
    delete p;
    return rc;
  }
  const long long rows = (long long)d->lin_batch * d->lin_channels;
  const long long total = rows * p->nr;
  double* dL = nullptr;
  if (hipMalloc(&dL, sizeof(double) * total * (cplx ? 2 : 1)) != hipSuccess ||
      hipMalloc(&p->tables, sizeof(float) * total * p->ncoef) != hipSuccess) {
    hipFree(dL);
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_ALLOC, "table allocation failed (%lld modes)", total);
  }
  double* dF = nullptr;
  rc = upload_nonlin_factor(d, rows, &dF);
  if (rc) {
    hipFree(dL);
    exmi_plan_destroy(p);
    return rc;
  }
  hipError_t e = hipMemcpy(dL, d->host_linear_operator, sizeof(double) * total, hipMemcpyHostToDevice);
  if (e == hipSuccess && cplx)
    e = hipMemcpy(dL + total, d->host_linear_operator_imag, sizeof(double) * total, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_coef(dL, cplx ? dL + total : nullptr, total, p->nr, d->dt, d->order, d->num_circle_points,
                    d->circle_radius, dF, p->tables, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  hipFree(dL);
  hipFree(dF);
  if (e != hipSuccess) {
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_HIP, "coefficient formation failed: %s", hipGetErrorString(e));
  }
  if (hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&p->launch_stream, hipStreamNonBlocking) != hipSuccess) {
    exmi_plan_destroy(p);
    return fail(EXMI_ERR_HIP, "stream creation failed");
  }
  if (!p->vjp) keep_for_twin(p, d, total);  // the multi-pass adjoint, on first use
  *out = p;
  return EXMI_OK;
}

int exmi_plan_destroy(exmi_plan* p) {
  if (!p) return EXMI_OK;
  {
    std::lock_guard<std::mutex> lk(g_group_mu);
    for (auto it = g_group_graphs.begin(); it != g_group_graphs.end();) {
      bool uses = it->first == p;
      for (size_t i = 0; i < it->second.key.size() && !uses; i += 4) uses = it->second.key[i] == p;
      if (uses) {
        if (it->second.exec) hipGraphExecDestroy(it->second.exec);
        it = g_group_graphs.erase(it);
      } else {
        ++it;
      }
    }
  }
  if (p->gc.exec) hipGraphExecDestroy(p->gc.exec);
  if (p->capture_stream) hipStreamDestroy(p->capture_stream);
  if (p->launch_stream) hipStreamDestroy(p->launch_stream);
  hipFree(p->tables);
  hipFree(p->work);
  hipFree(p->W3);
  hipFree(p->K3);
  hipFree(p->P3);
  hipFree(p->N03);
  hipFree(p->NC3);
  hipFree(p->S3);
  hipFree(p->S13);
  hipFree(p->A3);
  hipFree(p->F3);
  hipFree(p->R3);
  hipFree(p->PR3);
  hipFree(p->S0);
  hipFree(p->S1);
  hipFree(p->N0);
  hipFree(p->F);
  hipFree(p->T2);
  hipFree(p->ACC);
  hipFree(p->G);
  hipFree(p->LAM);
  hipFree(p->RHO);
  hipFree(p->H);
  hipFree(p->P);
  hipFree(p->ckpt);
  for (cf* b : {p->XA3, p->XB3, p->XC3, p->LAM3, p->MU3, p->T3, p->LA3, p->LB3, p->LC3, p->FH3}) hipFree(b);
  hipFree(p->RH3);
  hipFree(p->Q3);
  hipFree(p->coop_sync);
  hipFree(p->Gseg);
  if (p->twin) exmi_plan_destroy(p->twin);
  delete p;
  return EXMI_OK;
}

// Strategies a plan can roll out with: -1 per-step launches (not for plans
// forced to keep the spectrum on chip), 1 and 2 the multi-step kernels (2-D:
// 1 = the workgroup-resident kernel).
static bool strategy_ok(const exmi_plan* p, int nt) {
  if (nt == -1) return !(p->fourier_resident && !p->fr_auto);
  if (p->md) return nt == 2 && p->md_carry;
  if (p->ndim == 2) return (nt == 1 && p->roll2d != nullptr) || (nt == 2 && p->fr2d) || (nt == 3 && p->coop2d);
  return (nt == 1 || nt == 2) && p->roll[nt] != nullptr;
}

int exmi_plan_rollout_strategy(const exmi_plan* p) { return p ? p->roll_nt : fail(EXMI_ERR_INVALID, "null plan"); }

int exmi_plan_fft_config(const exmi_plan* p, int32_t* info, int32_t len) {
  if (!p || (!info && len > 0)) return fail(EXMI_ERR_INVALID, "null argument");
  if (!p->gen) return 0;
  int32_t v[4 + GFFT_MAX_PASS];
  int n = 0;
  v[n++] = p->gcode;
  v[n++] = p->gtuned;
  v[n++] = p->gcfg.lpb;
  v[n++] = p->gcfg.npass;
  for (int i = 0; i < p->gcfg.npass; ++i) v[n++] = p->gcfg.radix[i];
  for (int i = 0; i < n && i < len; ++i) info[i] = v[i];
  return n;
}

int exmi_plan_set_fft_config(exmi_plan* p, int32_t code) {
  if (!p) return fail(EXMI_ERR_INVALID, "null plan");
  if (!p->gen) return fail(EXMI_ERR_UNSUPPORTED, "the plan's transforms are the fused kernels' (no knobs)");
  GfftCfg c;
  if (!gfft_cfg_from_code(p->d.num_points, code, &c))
    return fail(EXMI_ERR_INVALID, "transform knobs %d: max_radix * 100000 + points_per_workgroup expected", code);
  p->gcfg = c;
  p->gcode = code;
  p->gtuned = 3;
  if (p->gc.exec) {
    hipGraphExecDestroy(p->gc.exec);
    p->gc = GraphCache();
  }
  return EXMI_OK;
}

int exmi_plan_set_rollout_strategy(exmi_plan* p, int32_t nt) {
  if (!p) return fail(EXMI_ERR_INVALID, "null plan");
  if (nt == -1 && p->fourier_resident && !p->fr_auto)
    return fail(EXMI_ERR_INVALID, "Fourier-resident plans roll out with the multi-step kernel only");
  if (!(nt == 0 || strategy_ok(p, nt)))
    return fail(EXMI_ERR_UNSUPPORTED, "rollout strategy %d not available for this plan", nt);
  if (nt == 0 && ((p->md && !p->md_carry) || (!p->md && p->ndim != 1 && !p->roll2d && !p->fr2d)))
    nt = -1;  // nothing to choose from
  p->roll_nt = nt;
  p->tune_short = 0;
  if (p->gc.exec) {
    hipGraphExecDestroy(p->gc.exec);
    p->gc = GraphCache();
  }
  return EXMI_OK;
}

int exmi_plan_num_coefs(const exmi_plan* p) { return p ? p->ncoef : fail(EXMI_ERR_INVALID, "null plan"); }

int exmi_plan_get_tables(const exmi_plan* p, float* host_out, int64_t host_count) {
  if (!p || !host_out) return fail(EXMI_ERR_INVALID, "null argument");
  const long long n = (long long)p->d.lin_batch * p->d.lin_channels * p->ncoef * p->nr;
  if (host_count < n) return fail(EXMI_ERR_INVALID, "host buffer too small: need %lld floats", n);
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(host_out, p->tables, sizeof(float) * n, hipMemcpyDeviceToHost));
  return EXMI_OK;
}

static int enqueue_roll(exmi_plan* p, int nt, const float* u0, long long u0_bs, float* out, long long out_bs,
                        long long out_ts, int n, int final_only, hipStream_t st, int* first_bad = nullptr,
                        float bound = 0.f, int bad_offset = 0);

static int run_step(exmi_plan* p, const void* in, long long in_bs, int in_f, void* out, long long out_bs, int out_f,
                    int nonlin_only, hipStream_t st) {
  Step1DArgs a;
  fill_args(p, &a);
  a.in = in;
  a.out = out;
  a.in_bstride = in_bs;
  a.out_bstride = out_bs;
  a.in_fourier = in_f;
  a.out_fourier = out_f;
  a.nonlin_only = nonlin_only;
  if (p->coupled) {
    if (nonlin_only || in_f || out_f)
      return fail(EXMI_ERR_UNSUPPORTED, "channel-coupled reaction plans step physical states only");
    const long long sz = (long long)p->d.num_channels * p->d.num_points;
    return enqueue_roll(p, 1, (const float*)in, in_bs, (float*)out, out_bs, sz, 1, 0, st);
  }
  hipError_t e = nonlin_only ? p->nonlin(a, st) : p->step(a, st);
  if (e != hipSuccess) return fail(EXMI_ERR_HIP, "step launch failed: %s", hipGetErrorString(e));
  return EXMI_OK;
}

int exmi_step(exmi_plan* p, const float* u_in, int64_t in_bs, float* u_out, int64_t out_bs, void* stream) {
  if (!p || !u_in || !u_out) return fail(EXMI_ERR_INVALID, "null argument");
  if (p->md) return enqueue_step3d(p, u_in, in_bs, u_out, out_bs, (hipStream_t)stream);
  if (p->ndim == 2) return enqueue_step2d(p, u_in, in_bs, u_out, out_bs, false, false, (hipStream_t)stream);
  return run_step(p, u_in, in_bs, 0, u_out, out_bs, 0, 0, (hipStream_t)stream);
}

int exmi_step_fourier(exmi_plan* p, const float* uh_in, int64_t in_bs, float* uh_out, int64_t out_bs,
                      void* stream) {
  if (!p || !uh_in || !uh_out) return fail(EXMI_ERR_INVALID, "null argument");
  if (p->md) return enqueue_step3d(p, uh_in, in_bs, uh_out, out_bs, (hipStream_t)stream, true, true);
  if (p->ndim == 2) return step_fourier2d(p, (const cf*)uh_in, in_bs, (cf*)uh_out, out_bs, (hipStream_t)stream);
  return run_step(p, uh_in, in_bs, 1, uh_out, out_bs, 1, 0, (hipStream_t)stream);
}

int exmi_nonlinear(exmi_plan* p, const float* uh_in, int64_t in_bs, float* nh_out, int64_t out_bs, void* stream) {
  if (!p || !uh_in || !nh_out) return fail(EXMI_ERR_INVALID, "null argument");
  if (p->d.order == 0) return fail(EXMI_ERR_INVALID, "order-0 plan has no nonlinear function");
  if (p->md) return enqueue_step3d(p, uh_in, in_bs, nh_out, out_bs, (hipStream_t)stream, true, true, true);
  if (p->ndim == 2) {
    // the nonlinear evaluation of the 2-D step chain, from and to spectra
    // (no exponential tables are read, so real and complex L plans share it)
    const int N = p->d.num_points;
    const long long szs = (long long)p->d.num_channels * N * (N / 2 + 1);
    if (in_bs != szs || out_bs != szs)
      return fail(EXMI_ERR_UNSUPPORTED, "exmi_nonlinear: 2-D spectra must be contiguous");
    Step2DArgs a;
    fill_args2d(p, &a);
    a.spec = (cf*)const_cast<float*>(uh_in);
    a.spec_out = (cf*)nh_out;
    a.spec_bstride = (long long)N * (N / 2 + 1);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = p->col2d[COL_CA_SPEC](a, p->ninst, st);
    if (e == hipSuccess) e = p->row2d[ROW_RB](a, p->ninst, st);
    if (e == hipSuccess) e = p->col2d[COL_NOUT](a, p->ninst, st);
    return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D nonlinear launch failed: %s", hipGetErrorString(e));
  }
  return run_step(p, uh_in, in_bs, 1, nh_out, out_bs, 1, 1, (hipStream_t)stream);
}

// strided batch copy: dst[b*ds + i] = src[b*ss + i], i < n
__global__ void exmi_copy_kernel(const float* __restrict__ src, long long ss, float* __restrict__ dst, long long ds,
                                 long long n, int B) {
  for (int b = blockIdx.y; b < B; b += gridDim.y)
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
      dst[(long long)b * ds + i] = src[(long long)b * ss + i];
}

static hipError_t copy_batch(const float* src, long long ss, float* dst, long long ds, long long n, int B,
                             hipStream_t st) {
  int gx = (int)((n + 255) / 256);
  if (gx > 64) gx = 64;
  const int gy = B < 65535 ? B : 65535;
  hipLaunchKernelGGL(exmi_copy_kernel, dim3(gx, gy), dim3(256), 0, st, src, ss, dst, ds, n, B);
  return hipGetLastError();
}

static int enqueue_roll(exmi_plan* p, int nt, const float* u0, long long u0_bs, float* out, long long out_bs,
                        long long out_ts, int n, int final_only, hipStream_t st, int* first_bad,
                        float bound, int bad_offset) {
  Roll1DArgs ra;
  fill_args(p, &ra.s);
  ra.first_bad = first_bad;
  ra.bound = bound;
  ra.bad_offset = bad_offset;
  ra.s.in = u0;
  ra.s.in_bstride = u0_bs;
  ra.traj = out;
  ra.traj_bstride = out_bs;
  ra.traj_tstride = out_ts;
  ra.nsteps = n;
  ra.final_only = final_only;
  hipError_t e = p->roll[nt](ra, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "rollout launch failed: %s", hipGetErrorString(e));
}

// 2-D workgroup-resident rollout: one launch for all n steps, instance i's
// states to out + i * out_bs + t * out_ts (only the last with final_only).
static int enqueue_roll2d(exmi_plan* p, const float* u0, long long u0_bs, float* out, long long out_bs,
                          long long out_ts, int n, int final_only, hipStream_t st) {
  Roll2DArgs ra;
  fill_args2d(p, &ra.s);
  ra.u0 = u0;
  ra.u0_bstride = u0_bs;
  ra.traj = out;
  ra.traj_bstride = out_bs;
  ra.traj_tstride = out_ts;
  ra.nsteps = n;
  ra.final_only = final_only;
  ra.ninst = p->ninst;
  const hipError_t e = p->roll2d(ra, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D rollout launch failed: %s", hipGetErrorString(e));
}

// 2-D spectrum-resident launch chain: RA CA once, then per step
// RB|RCB CB RB CCA, and a final RC. RCB emits the previous step's state
// while forming the products; CCA keeps S' as the next S0 with its fields.
// States go to out + i * out_bs + t * out_ts (only the last, at out, with
// final_only).
static int enqueue_chain_fr2d(exmi_plan* p, const float* u0, long long u0_bs, float* out, long long out_bs,
                              long long out_ts, int n, int final_only, hipStream_t st) {
  Step2DArgs a;
  fill_args2d(p, &a);
  a.u_in = u0;
  a.u_in_bstride = u0_bs;
  a.u_out_bstride = out_bs;
  a.write_T = 0;
  const int B = p->ninst;
  hipError_t e = p->row2d[ROW_RA](a, B, st);
  if (e == hipSuccess) e = p->col2d[COL_CA](a, B, st);
  for (int t = 0; t < n && e == hipSuccess; ++t) {
    if (t == 0 || final_only) {
      e = p->row2d[ROW_RB](a, B, st);
    } else {
      a.u_out = out + (long long)(t - 1) * out_ts;
      e = p->row2d[ROW_RCB](a, B, st);
    }
    if (e == hipSuccess) e = p->col2d[COL_CB](a, B, st);
    if (e == hipSuccess) e = p->row2d[ROW_RB](a, B, st);
    a.next_fields = t < n - 1 ? 1 : 0;
    if (e == hipSuccess) e = p->col2d[COL_CCA](a, B, st);
  }
  if (e == hipSuccess) {
    a.u_out = final_only ? out : out + (long long)(n - 1) * out_ts;
    e = p->row2d[ROW_RC](a, B, st);
  }
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D rollout launch failed: %s", hipGetErrorString(e));
}

// The chain as one cooperative launch for all n steps (states as in
// enqueue_chain_fr2d).
static int enqueue_coop2d(exmi_plan* p, const float* u0, long long u0_bs, float* out, long long out_bs,
                          long long out_ts, int n, int final_only, hipStream_t st, float* G = nullptr,
                          long long G_tstride = 0) {
  Coop2DArgs ca;
  memset(&ca, 0, sizeof ca);
  fill_args2d(p, &ca.s);
  ca.G = G;
  ca.G_tstride = G_tstride;
  ca.u0 = u0;
  ca.u0_bstride = u0_bs;
  ca.traj = out;
  ca.traj_bstride = out_bs;
  ca.traj_tstride = out_ts;
  ca.nsteps = n;
  ca.final_only = final_only;
  ca.ninst = p->ninst;
  ca.sync = p->coop_sync;
  ca.spin_limit = p->coop_spin;
  const hipError_t e = p->coop2d(ca, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D cooperative rollout launch failed: %s", hipGetErrorString(e));
}

// Rollout as per-step launches (nt <= 0) or one multi-step launch carrying
// nt transforms per lane group.
static int enqueue_rollout_nt(exmi_plan* p, int nt, const float* u0, long long u0_bs, float* traj, int n,
                              int include_init, hipStream_t st) {
  const long long sz = state_size(p);
  const long long tstride = (long long)(n + (include_init ? 1 : 0)) * sz;
  int off = 0;
  if (include_init) {
    hipError_t e = copy_batch(u0, u0_bs, traj, tstride, sz, p->d.batch, st);
    if (e != hipSuccess) return fail(EXMI_ERR_HIP, "copy launch failed: %s", hipGetErrorString(e));
    off = 1;
  }
  if (n == 0) return EXMI_OK;
  if (p->md) {
    if (nt == 2 && p->md_carry) return enqueue_md_carry(p, u0, u0_bs, traj + off * sz, tstride, sz, n, 0, st);
    for (int t = 0; t < n; ++t) {
      const float* in = (t == 0 && !include_init) ? u0 : traj + (long long)(t - 1 + off) * sz;
      const long long ibs = (t == 0 && !include_init) ? u0_bs : tstride;
      int rc = enqueue_step3d(p, in, ibs, traj + (long long)(t + off) * sz, tstride, st);
      if (rc) return rc;
    }
    return EXMI_OK;
  }
  if (p->ndim == 2) {
    if (nt == 1 && p->roll2d) return enqueue_roll2d(p, u0, u0_bs, traj + off * sz, tstride, sz, n, 0, st);
    if (nt == 2 && p->fr2d) return enqueue_chain_fr2d(p, u0, u0_bs, traj + off * sz, tstride, sz, n, 0, st);
    if (nt == 3 && p->coop2d) return enqueue_coop2d(p, u0, u0_bs, traj + off * sz, tstride, sz, n, 0, st);
    for (int t = 0; t < n; ++t) {
      int rc = enqueue_step2d(p, u0, u0_bs, traj + (long long)(t + off) * sz, tstride, t > 0, t < n - 1, st);
      if (rc) return rc;
    }
    return EXMI_OK;
  }
  // nt = -1 on a plan that chose the Fourier-resident kernels itself: the
  // per-step launches (the literal physical round trip) stay available
  if (p->fourier_resident && (nt == 0 || (nt < 0 && !p->fr_auto))) nt = p->roll[2] ? 2 : 1;
  if (nt > 0 && p->roll[nt]) return enqueue_roll(p, nt, u0, u0_bs, traj + off * sz, tstride, sz, n, 0, st);
  for (int t = 0; t < n; ++t) {
    const float* in = (t == 0 && !include_init) ? u0 : traj + (long long)(t - 1 + off) * sz;
    const long long ibs = (t == 0 && !include_init) ? u0_bs : tstride;
    float* out = traj + (long long)(t + off) * sz;
    int rc = run_step(p, in, ibs, 0, out, tstride, 0, 0, st);
    if (rc) return rc;
  }
  return EXMI_OK;
}

// Autotuner: time each available rollout strategy on a short probe into the
// caller's own buffer (overwritten by the real run) and keep the fastest.
// Autotuner cache: with EXMI_TUNE_CACHE naming a file, chosen strategies
// persist across processes as plain-text lines "<key> <strategy>", keyed by
// GPU architecture and the plan's shape (D, N, C, batch, kernel kind, order,
// per-trajectory tables, Fourier-resident).
static std::mutex g_tune_mu;

// Rollout strategies are timed on a probe of kTuneProbe steps (fewer when the
// first rollout is shorter; such a choice is provisional and timed again on
// the first rollout of at least kTuneProbe steps).
static constexpr int kTuneProbe = 16;
static bool needs_tune(const exmi_plan* p, int n) { return p->roll_nt == 0 || (p->tune_short && n >= kTuneProbe); }

static std::string tune_key(const exmi_plan* p) {
  hipDeviceProp_t prop;
  char arch[64] = "unknown";
  if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) {
    snprintf(arch, sizeof arch, "%s", prop.gcnArchName);
    for (char* c = arch; *c; ++c)
      if (*c == ' ' || *c == ':') *c = '_';
  }
  char key[256];
  snprintf(key, sizeof key, "%s|D%d|N%d|C%d|B%d|k%d|o%d|lb%d|fr%d", arch, p->ndim, p->d.num_points,
           p->d.num_channels, p->d.batch, p->kind, p->d.order, p->d.lin_batch,
           p->fourier_resident ? (p->fr_auto ? 2 : 1) : 0);
  return key;
}

static bool tune_lookup(exmi_plan* p) {
  const char* path = getenv("EXMI_TUNE_CACHE");
  if (!path || !*path) return false;
  const std::string key = tune_key(p);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char k[256];
  int v = 0;
  bool found = false;
  while (fscanf(f, "%255s %d", k, &v) == 2) {  // the last entry for a key wins
    if (key == k && strategy_ok(p, v)) {
      p->roll_nt = v;
      found = true;
    }
  }
  fclose(f);
  return found;
}

static void tune_store(const exmi_plan* p) {
  const char* path = getenv("EXMI_TUNE_CACHE");
  if (!path || !*path) return;
  const std::string key = tune_key(p);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (FILE* f = fopen(path, "a")) {
    fprintf(f, "%s %d\n", key.c_str(), p->roll_nt);
    fclose(f);
  }
}

// Runtime-size transforms: the radices (largest pass max_radix) and the
// lines per workgroup are timed at plan creation on a forward and an
// inverse transform of the plan's own batch, and the fastest kept; with
// EXMI_TUNE_CACHE the choice is looked up and stored like the rollout
// strategies ("<arch>|gfft|D|N|count <code>").
static std::string gfft_tune_key(const exmi_plan* p) {
  std::string k = tune_key(p);
  char key[256];
  snprintf(key, sizeof key, "%s|gfft|D%d|N%d|I%d", k.substr(0, k.find('|')).c_str(), p->ndim, p->d.num_points,
           p->ninst);
  return key;
}

static bool same_cfg(const GfftCfg& a, const GfftCfg& b) {
  if (a.n != b.n || a.npass != b.npass || a.lpb != b.lpb) return false;
  bool big = false;
  for (int i = 0; i < a.npass; ++i) big = big || (a.radix[i] > 8 && (a.radix[i] <= 17 || a.radix[i] == 19));
  if (big && a.reg19 != b.reg19) return false;
  for (int i = 0; i < a.npass; ++i)
    if (a.radix[i] != b.radix[i]) return false;
  return true;
}

static int tune_gfft(exmi_plan* p) {
  const int N = p->d.num_points;
  const std::string key = gfft_tune_key(p);
  const char* path = getenv("EXMI_TUNE_CACHE");
  if (path && *path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    bool found = false;
    if (FILE* f = fopen(path, "r")) {
      char k[256];
      int v = 0;
      GfftCfg c;
      while (fscanf(f, "%255s %d", k, &v) == 2) {  // the last entry for a key wins
        if (key == k && gfft_cfg_from_code(N, v, &c)) {
          p->gcode = v;
          p->gcfg = c;
          found = true;
        }
      }
      fclose(f);
    }
    if (found) {
      p->gtuned = 2;
      return EXMI_OK;
    }
  }
  // probe on up to ~4M points of the plan's batch, in its own buffers
  long long I = (1LL << 22) / p->n3;
  if (I < 1) I = 1;
  if (I > p->ninst) I = p->ninst;
  hipStream_t st = p->capture_stream;
  HIP_TRY(hipMemsetAsync(p->P3, 0, sizeof(float) * I * p->n3, st));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  std::vector<GfftCfg> seen;
  float best = 1e30f;
  int best_code = p->gcode;
  hipError_t err = hipSuccess;
  for (int i = 0; i < gfft_num_candidates() && err == hipSuccess; ++i) {
    const int code = gfft_candidate(i);
    GfftCfg c;
    if (!gfft_cfg_from_code(N, code, &c)) continue;
    bool dup = false;
    for (const GfftCfg& o : seen) dup = dup || same_cfg(o, c);
    if (dup) continue;
    seen.push_back(c);
    float tmin = 1e30f;
    for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {  // the first run of each candidate is not timed
      err = hipEventRecord(e0, st);
      if (err == hipSuccess) err = rfft_gen(c, p->ndim, I, p->P3, p->W3, p->tw_col.tw, st);
      if (err == hipSuccess) err = irfft_gen(c, p->ndim, I, p->W3, p->P3, p->tw_col.tw, st);
      if (err == hipSuccess) err = hipEventRecord(e1, st);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0 && ms < tmin) tmin = ms;
    }
    if (tmin < best) {
      best = tmin;
      best_code = code;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (err != hipSuccess) return fail(EXMI_ERR_HIP, "transform autotuning failed: %s", hipGetErrorString(err));
  gfft_cfg_from_code(N, best_code, &p->gcfg);
  p->gcode = best_code;
  p->gtuned = 1;
  if (path && *path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (FILE* f = fopen(path, "a")) {
      fprintf(f, "%s %d\n", key.c_str(), p->gcode);
      fclose(f);
    }
  }
  return EXMI_OK;
}

static int autotune_rollout(exmi_plan* p, const float* u0, long long u0_bs, float* traj, int n, int include_init) {
  if (!needs_tune(p, n)) return EXMI_OK;
  if (p->roll_nt == 0 && tune_lookup(p)) return EXMI_OK;
  const int probe = n < kTuneProbe ? n : kTuneProbe;
  p->tune_short = probe < kTuneProbe;
  if (probe == 0) {
    p->roll_nt = strategy_ok(p, 2) ? 2 : strategy_ok(p, 1) ? 1 : -1;
    return EXMI_OK;
  }
  hipStream_t st = p->capture_stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  float best = 1e30f;
  int best_nt = -1;
  const int cands[4] = {-1, 1, 2, 3};
  for (int ci = 0; ci < 4; ++ci) {
    const int nt = cands[ci];
    if (!strategy_ok(p, nt)) continue;
    float tmin = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {  // the first launch of each candidate is not timed
      HIP_TRY(hipEventRecord(e0, st));
      int rc = enqueue_rollout_nt(p, nt, u0, u0_bs, traj, probe, include_init, st);
      if (rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
      }
      HIP_TRY(hipEventRecord(e1, st));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms < tmin) tmin = ms;
    }
    if (tmin < best) {
      best = tmin;
      best_nt = nt;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  p->roll_nt = best_nt;
  if (!p->tune_short) tune_store(p);
  return EXMI_OK;
}

static int enqueue_rollout(exmi_plan* p, const float* u0, long long u0_bs, float* traj, int n, int include_init,
                           hipStream_t st) {
  return enqueue_rollout_nt(p, p->roll_nt, u0, u0_bs, traj, n, include_init, st);
}

static int run_maybe_graph(exmi_plan* p, const void* a, const void* b, long long s, int n, int flag, int kind,
                           int use_graph, hipStream_t st, int (*enq)(exmi_plan*, const void*, const void*, long long,
                                                                    int, int, hipStream_t),
                           bool launch = true) {
  if (!use_graph) return enq(p, a, b, s, n, flag, st);
  GraphCache& g = p->gc;
  if (!(g.exec && g.a == a && g.b == b && g.s == s && g.n == n && g.flag == flag && g.kind == kind)) {
    if (g.exec) {
      hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
    int rc = enq(p, a, b, s, n, flag, p->capture_stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(p->capture_stream, &graph);
    if (rc) {
      if (graph) hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(EXMI_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g.exec = nullptr;
      return fail(EXMI_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    g.a = a;
    g.b = b;
    g.s = s;
    g.n = n;
    g.flag = flag;
    g.kind = kind;
  }
  if (launch) HIP_TRY(hipGraphLaunch(g.exec, st));
  return EXMI_OK;
}

static int enq_rollout_cb(exmi_plan* p, const void* a, const void* b, long long s, int n, int flag, hipStream_t st) {
  return enqueue_rollout(p, (const float*)a, s, (float*)b, n, flag, st);
}

// The strategy a plan runs with before the autotuner has chosen one (repeat
// and the rollout gradient, which do not probe): the multi-step kernels
// where the plan has them, by the measurements behind the autotuner's
// usual picks (2-D: the workgroup-resident kernel up to 64^2, then the
// cooperative launch, then the launch chain).
static int effective_nt(const exmi_plan* p) {
  if (p->roll_nt != 0) return p->roll_nt;
  if (p->md) return p->md_carry ? 2 : -1;
  if (p->ndim == 2) {
    if (p->roll2d && p->d.num_points <= 64) return 1;
    if (p->coop2d) return 3;
    return p->fr2d ? 2 : -1;
  }
  return p->roll[2] ? 2 : (p->roll[1] ? 1 : -1);
}

static int enq_repeat_cb(exmi_plan* p, const void* a, const void* b, long long s, int n, int, hipStream_t st) {
  const float* u0 = (const float*)a;
  float* out = (float*)b;
  const long long sz = state_size(p);
  if (n == 0) {
    hipError_t e = copy_batch(u0, s, out, sz, sz, p->d.batch, st);
    return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
  }
  if (p->md) {
    if (effective_nt(p) == 2 && p->md_carry) return enqueue_md_carry(p, u0, s, out, sz, 0, n, 1, st);
    for (int t = 0; t < n; ++t) {
      int rc = enqueue_step3d(p, t == 0 ? u0 : out, t == 0 ? s : sz, out, sz, st);
      if (rc) return rc;
    }
    return EXMI_OK;
  }
  if (p->ndim == 2) {
    const int nt2 = effective_nt(p);
    if (nt2 == 1 && p->roll2d) return enqueue_roll2d(p, u0, s, out, sz, 0, n, 1, st);
    if (nt2 == 2 && p->fr2d) return enqueue_chain_fr2d(p, u0, s, out, sz, 0, n, 1, st);
    if (nt2 == 3 && p->coop2d) return enqueue_coop2d(p, u0, s, out, sz, 0, n, 1, st);
    for (int t = 0; t < n; ++t) {
      int rc = enqueue_step2d(p, u0, s, out, sz, t > 0, t < n - 1, st);
      if (rc) return rc;
    }
    return EXMI_OK;
  }
  const int nt = effective_nt(p);
  if (nt > 0 && p->roll[nt]) return enqueue_roll(p, nt, u0, s, out, sz, 0, n, 1, st);
  for (int t = 0; t < n; ++t) {
    const float* in = t == 0 ? u0 : (((n - t) % 2 == 0) ? out : p->work);
    const long long ibs = t == 0 ? s : sz;
    float* o = ((n - 1 - t) % 2 == 0) ? out : p->work;
    int rc = run_step(p, in, ibs, 0, o, sz, 0, 0, st);
    if (rc) return rc;
  }
  return EXMI_OK;
}

int exmi_rollout(exmi_plan* p, const float* u0, int64_t u0_bs, float* traj, int32_t n, int32_t include_init,
                 int32_t use_graph, void* stream) {
  if (!p || !u0 || !traj) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  if (needs_tune(p, n)) {
    // the probe writes into traj; order it after the caller's pending work
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    int rc = autotune_rollout(p, u0, u0_bs, traj, n, include_init);
    if (rc) return rc;
  }
  return run_maybe_graph(p, u0, traj, u0_bs, n, include_init ? 1 : 0, 1, use_graph, (hipStream_t)stream,
                         enq_rollout_cb);
}

int exmi_rollout_prepare(exmi_plan* p, const float* u0, int64_t u0_bs, float* traj, int32_t n,
                         int32_t include_init) {
  if (!p || !u0 || !traj) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  if (needs_tune(p, n)) {
    HIP_TRY(hipDeviceSynchronize());
    int rc = autotune_rollout(p, u0, u0_bs, traj, n, include_init);
    if (rc) return rc;
  }
  return run_maybe_graph(p, u0, traj, u0_bs, n, include_init ? 1 : 0, 1, 1, nullptr, enq_rollout_cb, false);
}

// ---- rollouts with failure detection (exmi_rollout_checked)
__global__ void exmi_fill_int_kernel(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void exmi_bad_finalize_kernel(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (p[i] == INT_MAX) p[i] = -1;
}
static unsigned small_grid(int n) { return (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024); }

}  // extern "C"
namespace exmi {
int launch_first_bad(int batch, int n_times, long long n, const float* x, long long bs, long long ts, float bound,
                     int offset, int* first_bad, hipStream_t st);
}
extern "C" {

int exmi_rollout_checked(exmi_plan* p, const float* u0, int64_t u0_bs, float* traj, int32_t n, int32_t include_init,
                         float bound, int32_t* first_bad, void* stream) {
  if (!p || !u0 || !traj || !first_bad) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  if (needs_tune(p, n)) {
    HIP_TRY(hipStreamSynchronize(st));
    int rc = autotune_rollout(p, u0, u0_bs, traj, n, include_init);
    if (rc) return rc;
  }
  const int B = p->d.batch;
  const float bnd = bound > 0.f ? bound : 3.402823466e38f;  // non-finite values fail every bound
  const long long sz = state_size(p);
  const int off = include_init ? 1 : 0;
  const long long tstride = (long long)(n + off) * sz;
  if (n == 0) {
    hipLaunchKernelGGL(exmi_fill_int_kernel, dim3(small_grid(B)), dim3(256), 0, st, (int*)first_bad, B, -1);
    HIP_TRY(hipGetLastError());
    return enqueue_rollout_nt(p, p->roll_nt, u0, u0_bs, traj, n, include_init, st);
  }
  int nt = p->roll_nt;
  if (p->ndim == 1 && !p->md && p->fourier_resident && (nt == 0 || (nt < 0 && !p->fr_auto)))
    nt = p->roll[2] ? 2 : 1;
  if (p->ndim == 1 && !p->md && nt > 0 && p->roll[nt]) {
    // the multi-step kernel checks every state it writes
    if (include_init) {
      hipError_t e = copy_batch(u0, u0_bs, traj, tstride, sz, B, st);
      if (e != hipSuccess) return fail(EXMI_ERR_HIP, "copy launch failed: %s", hipGetErrorString(e));
    }
    hipLaunchKernelGGL(exmi_fill_int_kernel, dim3(small_grid(B)), dim3(256), 0, st, (int*)first_bad, B, INT_MAX);
    HIP_TRY(hipGetLastError());
    int rc = enqueue_roll(p, nt, u0, u0_bs, traj + off * sz, tstride, sz, n, 0, st, (int*)first_bad, bnd, off);
    if (rc) return rc;
    hipLaunchKernelGGL(exmi_bad_finalize_kernel, dim3(small_grid(B)), dim3(256), 0, st, (int*)first_bad, B);
    HIP_TRY(hipGetLastError());
    return EXMI_OK;
  }
  // other strategies and dimensions: the rollout, then one scan of the
  // states it wrote
  int rc = enqueue_rollout_nt(p, p->roll_nt, u0, u0_bs, traj, n, include_init, st);
  if (rc) return rc;
  return launch_first_bad(B, n, sz, traj + off * sz, tstride, sz, bnd, off, (int*)first_bad, st);
}

// ---- grouped rollouts (see exmi_rollout_group)

// A member of a group of several plans rolls out with its tuned strategy,
// except that the cooperative launch (3) runs as the launch chain (2) it
// equals bit for bit: its workgroups stay resident for the whole rollout
// and, beside the other members' launches on concurrent streams, slowed
// the mixed Kolmogorov batch (32 x 128^2 + 16 x 256^2 + 8 x 512^2: 99.4
// against 86-92 us per step).
static int member_nt(const exmi_plan* p, int np) { return (np > 1 && p->roll_nt == 3) ? 2 : p->roll_nt; }

static int enqueue_member(exmi_plan* p, int np, const float* u0, long long u0_bs, float* traj, int n, int include_init,
                          hipStream_t st) {
  return enqueue_rollout_nt(p, member_nt(p, np), u0, u0_bs, traj, n, include_init, st);
}

static int enq_member_cb(exmi_plan* p, const void* a, const void* b, long long s, int n, int flag, hipStream_t st) {
  return enqueue_member(p, 2, (const float*)a, s, (float*)b, n, flag, st);
}

static int enqueue_group(exmi_plan* const* plans, int np, const float* const* u0s, const int64_t* bs,
                         float* const* trajs, int n, int include_init, hipStream_t main, bool fork) {
  if (!fork || np == 1) {
    for (int i = 0; i < np; ++i) {
      int rc = enqueue_member(plans[i], np, u0s[i], bs[i], trajs[i], n, include_init, main);
      if (rc) return rc;
    }
    return EXMI_OK;
  }
  std::vector<hipEvent_t> ev(np + 1);
  for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int rc = EXMI_OK;
  HIP_TRY(hipEventRecord(ev[0], main));
  for (int i = 1; i < np && !rc; ++i) {
    hipStream_t s = plans[i]->capture_stream;
    if (hipStreamWaitEvent(s, ev[0], 0) != hipSuccess) rc = fail(EXMI_ERR_HIP, "fork failed");
    if (!rc) rc = enqueue_member(plans[i], np, u0s[i], bs[i], trajs[i], n, include_init, s);
    if (!rc && hipEventRecord(ev[i], s) != hipSuccess) rc = fail(EXMI_ERR_HIP, "join record failed");
  }
  if (!rc) rc = enqueue_member(plans[0], np, u0s[0], bs[0], trajs[0], n, include_init, main);
  for (int i = 1; i < np && !rc; ++i)
    if (hipStreamWaitEvent(main, ev[i], 0) != hipSuccess) rc = fail(EXMI_ERR_HIP, "join failed");
  for (auto& e : ev) hipEventDestroy(e);
  return rc;
}

int exmi_rollout_group(exmi_plan* const* plans, int32_t np, const float* const* u0s, const int64_t* u0_bstrides,
                       float* const* trajs, int32_t n, int32_t include_init, int32_t use_graph, void* stream) {
  if (!plans || np < 1 || !u0s || !u0_bstrides || !trajs) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  for (int i = 0; i < np; ++i) {
    if (!plans[i] || !u0s[i] || !trajs[i]) return fail(EXMI_ERR_INVALID, "null plan or buffer in group");
    for (int k = 0; k < i; ++k)
      if (plans[k] == plans[i]) return fail(EXMI_ERR_INVALID, "a plan appears twice in the group");
    if (needs_tune(plans[i], n)) {
      HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
      int rc = autotune_rollout(plans[i], u0s[i], u0_bstrides[i], trajs[i], n, include_init);
      if (rc) return rc;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return enqueue_group(plans, np, u0s, u0_bstrides, trajs, n, include_init, st, false);
  if (use_graph == 2 || use_graph == 3) {
    // every plan's rollout runs on its own stream (a hardware queue of its
    // own), as a cached per-plan graph (2) or as direct launches (3), forked
    // from and joined back into `stream`
    if (use_graph == 2) {
      for (int i = 0; i < np; ++i) {
        // graph kind 3 (a group member's rollout) when its strategy differs
        // from the plan's own, so the plan's cached rollout graph is not reused
        const bool differs = member_nt(plans[i], np) != plans[i]->roll_nt;
        int rc = run_maybe_graph(plans[i], u0s[i], trajs[i], u0_bstrides[i], n, include_init ? 1 : 0, differs ? 3 : 1,
                                 1, nullptr, differs ? enq_member_cb : enq_rollout_cb, false);
        if (rc) return rc;
      }
    }
    std::vector<hipEvent_t> ev(np + 1);
    for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int rc = EXMI_OK;
    if (hipEventRecord(ev[0], st) != hipSuccess) rc = fail(EXMI_ERR_HIP, "fork record failed");
    for (int i = 0; i < np && !rc; ++i) {
      hipStream_t ls = plans[i]->launch_stream;
      if (hipStreamWaitEvent(ls, ev[0], 0) != hipSuccess) rc = fail(EXMI_ERR_HIP, "fork failed");
      if (!rc) {
        if (use_graph == 2) {
          if (hipGraphLaunch(plans[i]->gc.exec, ls) != hipSuccess) rc = fail(EXMI_ERR_HIP, "graph launch failed");
        } else {
          rc = enqueue_member(plans[i], np, u0s[i], u0_bstrides[i], trajs[i], n, include_init, ls);
        }
      }
      if (!rc && hipEventRecord(ev[i + 1], ls) != hipSuccess) rc = fail(EXMI_ERR_HIP, "join record failed");
    }
    for (int i = 0; i < np && !rc; ++i)
      if (hipStreamWaitEvent(st, ev[i + 1], 0) != hipSuccess) rc = fail(EXMI_ERR_HIP, "join failed");
    for (auto& e : ev) hipEventDestroy(e);
    return rc;
  }
  std::vector<const void*> key;
  for (int i = 0; i < np; ++i) {
    key.push_back(plans[i]);
    key.push_back(u0s[i]);
    key.push_back((const void*)(intptr_t)u0_bstrides[i]);
    key.push_back(trajs[i]);
  }
  std::lock_guard<std::mutex> lk(g_group_mu);
  GroupGraph& g = g_group_graphs[plans[0]];
  if (!(g.exec && g.key == key && g.n == n && g.flag == include_init)) {
    if (g.exec) {
      hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
    hipStream_t cs = plans[0]->capture_stream;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_group(plans, np, u0s, u0_bstrides, trajs, n, include_init, cs, true);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &graph);
    if (rc) {
      if (graph) hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(EXMI_ERR_HIP, "group graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g.exec = nullptr;
      return fail(EXMI_ERR_HIP, "group graph instantiation failed: %s", hipGetErrorString(e));
    }
    g.key = key;
    g.n = n;
    g.flag = include_init;
  }
  HIP_TRY(hipGraphLaunch(g.exec, st));
  return EXMI_OK;
}

int exmi_repeat(exmi_plan* p, const float* u0, int64_t u0_bs, float* u_out, int32_t n, int32_t use_graph,
                void* stream) {
  if (!p || !u0 || !u_out) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  if (n > 1) {
    int rc = ensure_work(p);
    if (rc) return rc;
  }
  return run_maybe_graph(p, u0, u_out, u0_bs, n, 0, 2, use_graph, (hipStream_t)stream, enq_repeat_cb);
}

}  // extern "C"

// 2-D transforms: the row kernel (r2c along y) and the column kernel (c2c
// along x) of the 2-D steppers, through a [ky][x] intermediate of
// (N/2+1) N complex entries per field: `T` when given (the caller's
// scratch), else a stream-ordered allocation. Twiddles `tr` (row size N/2)
// and `tc` (column size N).
static int check_2d_transform(int N, int count) {
  if (!pow2(N) || N < 32 || N > 512)
    return fail(EXMI_ERR_UNSUPPORTED, "2-D transforms: num_points=%d: supported sizes are powers of two 32..512", N);
  if (count > 65535) return fail(EXMI_ERR_UNSUPPORTED, "2-D transforms: at most 65535 fields per call");
  return EXMI_OK;
}

static int transform2d_with(bool inverse, int N, int count, const float* x, cf* xh, const Twiddles& tr,
                            const Twiddles& tc, cf* T, hipStream_t st) {
  int rc = check_2d_transform(N, count);
  if (rc) return rc;
  const Launch2D row = get_2d(N, NL_GRADNORM, 1, inverse ? ROW_RC : ROW_RA);
  const Launch2D col = get_2d(N, NL_GRADNORM, 0, inverse ? COL_INV : COL_FWD);
  if (!row || !col) return fail(EXMI_ERR_UNSUPPORTED, "no 2-D transform kernel for N=%d", N);
  const long long full = (long long)(N / 2 + 1) * N;
  Step2DArgs a;
  memset(&a, 0, sizeof a);
  a.tw_row = tr.tw;
  a.fw_row = tr.fw;
  a.tw_col = tc.tw;
  a.kmax = N;  // no dealiasing
  a.e_im = -1;
  a.channels = 1;  // one instance per field
  a.tstride = full;
  a.spec = xh;
  a.spec_bstride = full;
  if (T) a.T = T;
  else HIP_TRY(hipMallocAsync((void**)&a.T, sizeof(cf) * full * count, st));
  hipError_t e;
  if (!inverse) {
    a.u_in = x;
    a.u_in_bstride = (long long)N * N;
    e = row(a, count, st);
    if (e == hipSuccess) e = col(a, count, st);
  } else {
    a.u_out = const_cast<float*>(x);
    a.u_out_bstride = (long long)N * N;
    e = col(a, count, st);
    if (e == hipSuccess) e = row(a, count, st);
  }
  if (!T) {
    const hipError_t ef = hipFreeAsync(a.T, st);
    if (e == hipSuccess) e = ef;
  }
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "2-D transform launch failed: %s", hipGetErrorString(e));
}

static int transform2d(bool inverse, int N, int count, const float* x, cf* xh, hipStream_t st) {
  int rc = check_2d_transform(N, count);
  if (rc) return rc;
  Twiddles tr, tc;
  rc = get_twiddles(N / 2, &tr);
  if (!rc) rc = get_twiddles(N, &tc);
  if (rc) return rc;
  return transform2d_with(inverse, N, count, x, xh, tr, tc, nullptr, st);
}

// Standalone transforms of the sizes outside the fused kernels' set, with the
// runtime-size kernels and their default knobs; the inverse of D >= 2 runs
// its in-place line transforms on a stream-ordered copy of the spectra.
static int gen_transform(bool inverse, int D, int N, int count, const float* in, float* out, hipStream_t st) {
  if (!gen_size(N))
    return fail(EXMI_ERR_UNSUPPORTED, "%s: num_points=%d: supported sizes are 2..%d", inverse ? "exmi_irfft" : "exmi_rfft",
                N, GFFT_MAX_N);
  if (count <= 0) return EXMI_OK;
  if (!in || !out) return fail(EXMI_ERR_INVALID, "null argument");
  Twiddles tw;
  int rc = get_twiddles(N, &tw);
  if (rc) return rc;
  GfftCfg c;
  gfft_cfg_from_code(N, gfft_default_code(), &c);
  hipError_t e = hipSuccess;
  if (!inverse) {
    e = rfft_gen(c, D, count, in, (cf*)out, tw.tw, st);
  } else if (D == 1) {
    e = launch_gfft(c, GFFT_C2R, in, out, count, 1, 0, 1, 1.f / (float)N, tw.tw, st);
  } else {
    size_t nr = (size_t)(N / 2 + 1);
    for (int d = 1; d < D; ++d) nr *= (size_t)N;
    const size_t bytes = sizeof(cf) * nr * (size_t)count;
    cf* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, bytes, st));
    e = hipMemcpyAsync(tmp, in, bytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = irfft_gen(c, D, count, tmp, out, tw.tw, st);
    const hipError_t ef = hipFreeAsync(tmp, st);
    if (e == hipSuccess) e = ef;
  }
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "runtime-size transform launch failed: %s", hipGetErrorString(e));
}

extern "C" {

int exmi_rfft(int32_t D, int32_t N, int32_t count, const float* x, float* x_hat, void* stream) {
  if (D >= 1 && D <= 3 && !fused_size(D, N)) return gen_transform(false, D, N, count, x, x_hat, (hipStream_t)stream);
  if (D == 2) return count <= 0 ? EXMI_OK : transform2d(false, N, count, x, (cf*)x_hat, (hipStream_t)stream);
  if (D == 3) {
    if (!size_3d(N)) return fail(EXMI_ERR_UNSUPPORTED, "exmi_rfft: 3-D num_points=%d (powers of two 16..256)", N);
    if (count <= 0) return EXMI_OK;
    Twiddles tz, tl;
    int rc = get_twiddles(N / 2, &tz);
    if (!rc) rc = get_twiddles(N, &tl);
    if (rc) return rc;
    hipError_t e = rfft3(N, count, x, (cf*)x_hat, tz, tl, (hipStream_t)stream);
    return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "3-D rfft launch failed: %s", hipGetErrorString(e));
  }
  if (D != 1) return fail(EXMI_ERR_UNSUPPORTED, "exmi_rfft: num_spatial_dims=%d not supported", D);
  if (!size_1d(N)) return fail(EXMI_ERR_UNSUPPORTED, "exmi_rfft: num_points=%d", N);
  if (count <= 0) return EXMI_OK;
  Twiddles tw;
  int rc = get_twiddles(N / 2, &tw);
  if (rc) return rc;
  hipError_t e = launch_fft1d(N / 2, false, x, x_hat, count, tw.rt(), (hipStream_t)stream);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "rfft launch failed: %s", hipGetErrorString(e));
}

int exmi_poisson_fourier(int32_t D, int32_t N, int32_t count, double domain_extent, int32_t order, float* uh,
                         void* stream) {
  if (D < 1 || D > 3) return fail(EXMI_ERR_UNSUPPORTED, "exmi_poisson_fourier: num_spatial_dims 1, 2, 3");
  if (!uh) return fail(EXMI_ERR_INVALID, "null argument");
  if (N < 2 || count < 0 || !(domain_extent > 0) || order < 2 || order % 2)
    return fail(EXMI_ERR_INVALID, "exmi_poisson_fourier: sizes, or an odd order");
  if (count == 0) return EXMI_OK;
  const hipError_t e = launch_poisson_spec((cf*)uh, count, D, N, 2.0 * 3.14159265358979323846 / domain_extent, order,
                                           (hipStream_t)stream);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "Poisson launch failed: %s", hipGetErrorString(e));
}

int exmi_axpy(int64_t n, float a, const float* x, const float* y, float* out, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y || !out))) return fail(EXMI_ERR_INVALID, "exmi_axpy: arguments");
  if (n == 0) return EXMI_OK;
  const hipError_t e = launch_axpy(n, a, x, y, out, (hipStream_t)stream);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "axpy launch failed: %s", hipGetErrorString(e));
}

int exmi_wave_step_fourier(int32_t D, int32_t N, int32_t batch, double domain_extent, double speed_of_sound,
                           double dt, float* uh, void* stream) {
  if (D < 1 || D > 3) return fail(EXMI_ERR_UNSUPPORTED, "exmi_wave_step_fourier: num_spatial_dims 1, 2, 3");
  if (!uh) return fail(EXMI_ERR_INVALID, "null argument");
  if (N < 2 || batch < 0 || !(domain_extent > 0)) return fail(EXMI_ERR_INVALID, "exmi_wave_step_fourier: sizes");
  if (batch == 0) return EXMI_OK;
  const hipError_t e = launch_wave_spec((cf*)uh, batch, D, N, 2.0 * 3.14159265358979323846 / domain_extent,
                                        speed_of_sound, dt, (hipStream_t)stream);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "wave step launch failed: %s", hipGetErrorString(e));
}

int exmi_irfft(int32_t D, int32_t N, int32_t count, const float* x_hat, float* x, void* stream) {
  if (D >= 1 && D <= 3 && !fused_size(D, N)) return gen_transform(true, D, N, count, x_hat, x, (hipStream_t)stream);
  if (D == 2)
    return count <= 0 ? EXMI_OK : transform2d(true, N, count, x, (cf*)const_cast<float*>(x_hat), (hipStream_t)stream);
  if (D == 3) {
    if (!size_3d(N)) return fail(EXMI_ERR_UNSUPPORTED, "exmi_irfft: 3-D num_points=%d (powers of two 16..256)", N);
    if (count <= 0) return EXMI_OK;
    Twiddles tz, tl;
    int rc = get_twiddles(N / 2, &tz);
    if (!rc) rc = get_twiddles(N, &tl);
    if (rc) return rc;
    // the inverse line transforms run in place: on a stream-ordered copy
    const hipStream_t st = (hipStream_t)stream;
    const size_t bytes = sizeof(cf) * (size_t)count * N * N * (N / 2 + 1);
    cf* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, bytes, st));
    hipError_t e = hipMemcpyAsync(tmp, x_hat, bytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = irfft3(N, count, tmp, x, tz, tl, st);
    const hipError_t ef = hipFreeAsync(tmp, st);
    if (e == hipSuccess) e = ef;
    return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "3-D irfft launch failed: %s", hipGetErrorString(e));
  }
  if (D != 1) return fail(EXMI_ERR_UNSUPPORTED, "exmi_irfft: num_spatial_dims=%d not supported", D);
  if (!size_1d(N)) return fail(EXMI_ERR_UNSUPPORTED, "exmi_irfft: num_points=%d", N);
  if (count <= 0) return EXMI_OK;
  Twiddles tw;
  int rc = get_twiddles(N / 2, &tw);
  if (rc) return rc;
  hipError_t e = launch_fft1d(N / 2, true, x_hat, x, count, tw.rt(), (hipStream_t)stream);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "irfft launch failed: %s", hipGetErrorString(e));
}

}  // extern "C"

// ------------------------------------------- allocation-free inverse transform
// exmi_irfft_noalloc keeps every buffer it needs beyond the caller's input
// and output in two file-scope device arrays: `scratch` (the 2-D [ky][x]
// intermediate, or the copy of the 3-D spectra that the in-place line
// transforms work on) and `noalloc_twiddles` (the twiddle tables of every
// transform size, the same values get_twiddles forms). Each GPU holds its
// own copy of both until the process ends. In host code their names are
// stand-ins, not device addresses: the set-up takes each device's addresses
// with hipGetSymbolAddress, and only those are passed to kernels and copies.
static constexpr size_t kNoAllocScratchBytes = (size_t)EXMI_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static constexpr int kNoAllocSizes[] = {8, 16, 24, 25, 32, 64, 128, 256, 512, 1024, 2048};
static constexpr int noalloc_tw_total() {
  int t = 0;
  for (int m : kNoAllocSizes) t += 2 * m;
  return t;
}
static __device__ __align__(256) cf noalloc_twiddles[noalloc_tw_total()];

struct NoAllocDevice {
  unsigned char* scratch = nullptr;  // this device's `scratch`
  cf* tw = nullptr;                  // this device's `noalloc_twiddles`
};
static std::mutex g_noalloc_mu;
static std::map<int, NoAllocDevice> g_noalloc;  // by device

// Once per device, with that device current: the arrays' addresses and the
// twiddle values (a synchronous host-to-device copy, no allocation).
static int noalloc_setup(NoAllocDevice* out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  auto it = g_noalloc.find(dev);
  if (it != g_noalloc.end()) {
    *out = it->second;
    return EXMI_OK;
  }
  NoAllocDevice d;
  hipError_t e = hipGetSymbolAddress((void**)&d.scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return fail(EXMI_ERR_HIP, "scratch symbol address: %s", hipGetErrorString(e));
  e = hipGetSymbolAddress((void**)&d.tw, HIP_SYMBOL(noalloc_twiddles));
  if (e != hipSuccess) return fail(EXMI_ERR_HIP, "twiddle symbol address: %s", hipGetErrorString(e));
  std::vector<cf> h(noalloc_tw_total());
  int off = 0;
  for (int m : kNoAllocSizes) {
    twiddle_values(m, &h[off], &h[off + m]);
    off += 2 * m;
  }
  e = hipMemcpy(d.tw, h.data(), sizeof(cf) * h.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(EXMI_ERR_HIP, "twiddle upload: %s", hipGetErrorString(e));
  g_noalloc[dev] = d;
  *out = d;
  return EXMI_OK;
}

static bool noalloc_twiddles_of(const NoAllocDevice& d, int M, Twiddles* t) {
  int off = 0;
  for (int m : kNoAllocSizes) {
    if (m == M) {
      t->tw = d.tw + off;
      t->fw = d.tw + off + m;
      return true;
    }
    off += 2 * m;
  }
  return false;
}

extern "C" {

int exmi_irfft_noalloc(int32_t D, int32_t N, int32_t count, const float* x_hat, float* x, void* stream) {
  NoAllocDevice nd;
  int rc = noalloc_setup(&nd);
  if (rc) return rc;
  if (count < 0) return fail(EXMI_ERR_INVALID, "exmi_irfft_noalloc: count must be >= 0");
  if (D < 1 || D > 3) return fail(EXMI_ERR_UNSUPPORTED, "exmi_irfft_noalloc: num_spatial_dims=%d not supported", D);
  if (count == 0) return EXMI_OK;
  if (!x_hat || !x) return fail(EXMI_ERR_INVALID, "null argument");
  const hipStream_t st = (hipStream_t)stream;
  if (D == 1) {
    if (!size_1d(N)) return fail(EXMI_ERR_UNSUPPORTED, "exmi_irfft_noalloc: num_points=%d", N);
    Twiddles tw;
    if (!noalloc_twiddles_of(nd, N / 2, &tw)) return fail(EXMI_ERR_UNSUPPORTED, "no twiddles for N=%d", N);
    const hipError_t e = launch_fft1d(N / 2, true, x_hat, x, count, tw.rt(), st);
    return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "irfft launch failed: %s", hipGetErrorString(e));
  }
  if (D == 2) {
    rc = check_2d_transform(N, count);
    if (rc) return rc;
    const size_t need = sizeof(cf) * (size_t)(N / 2 + 1) * N * count;
    if (need > kNoAllocScratchBytes)
      return fail(EXMI_ERR_INVALID,
                  "exmi_irfft_noalloc: %d 2-D fields of N=%d need %zu bytes of scratch, more than its %zu bytes "
                  "(at most %zu fields per call)",
                  count, N, need, kNoAllocScratchBytes, kNoAllocScratchBytes / (sizeof(cf) * (N / 2 + 1) * N));
    Twiddles tr, tc;
    if (!noalloc_twiddles_of(nd, N / 2, &tr) || !noalloc_twiddles_of(nd, N, &tc))
      return fail(EXMI_ERR_UNSUPPORTED, "no twiddles for N=%d", N);
    return transform2d_with(true, N, count, x, (cf*)const_cast<float*>(x_hat), tr, tc, (cf*)nd.scratch, st);
  }
  if (!size_3d(N)) return fail(EXMI_ERR_UNSUPPORTED, "exmi_irfft_noalloc: 3-D num_points=%d (powers of two 16..256)", N);
  const size_t per = sizeof(cf) * (size_t)N * N * (N / 2 + 1);
  const size_t need = per * count;
  if (need > kNoAllocScratchBytes)
    return fail(EXMI_ERR_INVALID,
                "exmi_irfft_noalloc: %d 3-D fields of N=%d need %zu bytes of scratch, more than its %zu bytes "
                "(at most %zu fields per call)",
                count, N, need, kNoAllocScratchBytes, kNoAllocScratchBytes / per);
  Twiddles tz, tl;
  if (!noalloc_twiddles_of(nd, N / 2, &tz) || !noalloc_twiddles_of(nd, N, &tl))
    return fail(EXMI_ERR_UNSUPPORTED, "no twiddles for N=%d", N);
  // the inverse line transforms run in place: on a copy in the scratch
  cf* X = (cf*)nd.scratch;
  hipError_t e = copy_batch(x_hat, 0, (float*)X, 0, (long long)(need / sizeof(float)), 1, st);
  if (e == hipSuccess) e = irfft3(N, count, X, x, tz, tl, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "3-D irfft launch failed: %s", hipGetErrorString(e));
}

// ------------------------------------------------------- corrected rollout
}  // extern "C"
namespace exmi {
size_t cnn_frag_bytes();
hipError_t cnn_pack(const float* params, void* frag, hipStream_t st);
hipError_t cnn_forward_strided(const float* params, const void* frag, int N, long long count, const float* x,
                               long long xs, float* y, long long ys, hipStream_t st);
}  // namespace exmi
extern "C" {

int exmi_corrected_rollout(exmi_plan* p, const float* params, const float* u0, int64_t u0_bs, float* traj,
                           int32_t n, int32_t include_init, void* stream) {
  if (!p || !params || !u0 || !traj) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  const int N = p->d.num_points;
  if (p->ndim != 1 || p->d.num_channels != 1 || !(N == 32 || N == 64 || N == 128))
    return fail(EXMI_ERR_UNSUPPORTED, "corrected rollout: 1-D single-channel plans with 32, 64 or 128 points");
  hipStream_t st = (hipStream_t)stream;
  const int B = p->d.batch;
  const int off = include_init ? 1 : 0;
  const long long ts = (long long)(n + off) * N;  // trajectory batch stride
  if (include_init) {
    hipError_t e = copy_batch(u0, u0_bs, traj, ts, N, B, st);
    if (e != hipSuccess) return fail(EXMI_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
  }
  if (n == 0) return EXMI_OK;
  void* frag = nullptr;
  float* tmp = nullptr;
  if (hipMallocAsync(&frag, cnn_frag_bytes(), st) != hipSuccess)
    return fail(EXMI_ERR_ALLOC, "corrected rollout: workspace allocation failed");
  if (hipMallocAsync((void**)&tmp, sizeof(float) * (size_t)B * N, st) != hipSuccess) {
    (void)hipFreeAsync(frag, st);
    return fail(EXMI_ERR_ALLOC, "corrected rollout: workspace allocation failed");
  }
  int rc = EXMI_OK;
  hipError_t e = cnn_pack(params, frag, st);
  for (int t = 0; t < n && rc == EXMI_OK && e == hipSuccess; ++t) {
    const float* in = t == 0 ? u0 : traj + (long long)(t - 1 + off) * N;
    const long long ibs = t == 0 ? u0_bs : ts;
    rc = run_step(p, in, ibs, 0, tmp, N, 0, 0, st);
    if (rc == EXMI_OK) e = cnn_forward_strided(params, frag, N, B, tmp, N, traj + (long long)(t + off) * N, ts, st);
  }
  (void)hipFreeAsync(frag, st);
  (void)hipFreeAsync(tmp, st);
  if (rc) return rc;
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "corrected rollout launch failed: %s", hipGetErrorString(e));
}

// ------------------------------------------------------------------ adjoint
__global__ void exmi_axpy_kernel(float* __restrict__ dst, long long ds, const float* __restrict__ src, long long ss,
                                 long long n, int B) {
  for (int b = blockIdx.y; b < B; b += gridDim.y)
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
      dst[(long long)b * ds + i] += src[(long long)b * ss + i];
}

__global__ void exmi_fill_kernel(float* __restrict__ dst, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = v;
}

static hipError_t axpy_batch(float* dst, long long ds, const float* src, long long ss, long long n, int B,
                             hipStream_t st) {
  int gx = (int)((n + 255) / 256);
  if (gx > 64) gx = 64;
  const int gy = B < 65535 ? B : 65535;
  hipLaunchKernelGGL(exmi_axpy_kernel, dim3(gx, gy), dim3(256), 0, st, dst, ds, src, ss, n, B);
  return hipGetLastError();
}

static hipError_t fill(float* dst, long long n, float v, hipStream_t st) {
  long long gx = (n + 255) / 256;
  if (gx > 4096) gx = 4096;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(exmi_fill_kernel, dim3((unsigned)gx), dim3(256), 0, st, dst, n, v);
  return hipGetLastError();
}

static bool fused_adjoint(const exmi_plan* p) {
  return !p->md && (p->ndim == 1 ? p->vjp != nullptr : p->rowA[0] != nullptr);
}

static bool has_adjoint(const exmi_plan* p) {
  if (p->md) return md_adjoint_ok(p);
  return fused_adjoint(p) || !p->hL.empty();
}

static int ensure_adjoint_ws(exmi_plan* p) {
  if (p->md) return ensure_adj_md(p);
  if (!fused_adjoint(p)) return ensure_twin(p);
  if (p->ndim != 2 || p->G) return EXMI_OK;
  const int N = p->d.num_points;
  const long long B = p->d.batch, full = (long long)(N / 2 + 1) * N;
  bool ok = hipMalloc(&p->G, sizeof(float) * B * 4 * N * N) == hipSuccess &&
            hipMalloc(&p->LAM, sizeof(cf) * B * full) == hipSuccess &&
            hipMalloc(&p->RHO, sizeof(cf) * B * full) == hipSuccess &&
            hipMalloc(&p->H, sizeof(cf) * B * p->kc * N) == hipSuccess &&
            hipMalloc(&p->P, sizeof(cf) * B * 2 * p->kc * N) == hipSuccess;
  if (!ok) {
    hipFree(p->G);
    hipFree(p->LAM);
    hipFree(p->RHO);
    hipFree(p->H);
    hipFree(p->P);
    p->G = nullptr;
    p->LAM = p->RHO = p->H = p->P = nullptr;
    return fail(EXMI_ERR_ALLOC, "adjoint workspace allocation failed (batch %lld, N %d)", B, N);
  }
  return EXMI_OK;
}

// gout = (d step / d u)^T g at u; gout may alias g (g is read first).
static int enqueue_step_vjp(exmi_plan* p, const float* u, long long ubs, const float* g, long long gbs, float* gout,
                            long long obs, hipStream_t st) {
  if (p->md) return enqueue_vjp_md(p, u, ubs, g, gbs, gout, obs, st);
  if (!fused_adjoint(p)) {
    if (!p->twin) return fail(EXMI_ERR_UNSUPPORTED, "adjoint twin plan not built");
    return enqueue_vjp_md(p->twin, u, ubs, g, gbs, gout, obs, st);
  }
  const int B = p->d.batch;
  hipError_t e = hipSuccess;
  if (p->ndim == 1) {
    Vjp1DArgs va;
    fill_args(p, &va.s);
    va.s.in = u;
    va.s.in_bstride = ubs;
    va.g = g;
    va.g_bstride = gbs;
    va.gout = gout;
    va.gout_bstride = obs;
    e = p->vjp(va, st);
  } else {
    Adj2DArgs aa;
    memset(&aa, 0, sizeof aa);
    fill_args2d(p, &aa.s);
    aa.G = p->G;
    aa.LAM = p->LAM;
    aa.RHO = p->RHO;
    aa.H = p->H;
    aa.P = p->P;
    Step2DArgs& a = aa.s;
    // recompute the forward stage fields at u
    a.u_in = u;
    a.u_in_bstride = ubs;
    e = p->row2d[ROW_RA](a, B, st);
    if (e == hipSuccess) e = p->col2d[COL_CA](a, B, st);
    aa.stage = 0;
    if (e == hipSuccess) e = p->rowA[ROWA_RBG](aa, B, st);
    if (e == hipSuccess) e = p->col2d[COL_CB](a, B, st);
    aa.stage = 1;
    if (e == hipSuccess) e = p->rowA[ROWA_RG](aa, B, st);
    // transposed step
    a.u_in = g;
    a.u_in_bstride = gbs;
    if (e == hipSuccess) e = p->row2d[ROW_RA](a, B, st);
    if (e == hipSuccess) e = p->colA[COLA_CL](aa, B, st);
    aa.stage = 1;
    if (e == hipSuccess) e = p->rowA[ROWA_RP](aa, B, st);
    if (e == hipSuccess) e = p->colA[COLA_CD](aa, B, st);
    aa.stage = 0;
    if (e == hipSuccess) e = p->rowA[ROWA_RP](aa, B, st);
    if (e == hipSuccess) e = p->colA[COLA_CE](aa, B, st);
    a.u_out = gout;
    a.u_out_bstride = obs;
    a.write_T = 0;
    if (e == hipSuccess) e = p->row2d[ROW_RC](a, B, st);
  }
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "adjoint step launch failed: %s", hipGetErrorString(e));
}

static int ensure_work(exmi_plan* p) {
  const size_t need = (size_t)p->d.batch * p->d.num_channels * p->d.num_points;
  if (p->ndim != 1 || p->md || p->work_elems >= need) return EXMI_OK;
  hipFree(p->work);
  p->work = nullptr;
  p->work_elems = 0;
  if (hipMalloc(&p->work, need * sizeof(float)) != hipSuccess) return fail(EXMI_ERR_ALLOC, "workspace allocation failed");
  p->work_elems = need;
  return EXMI_OK;
}

// The plans the adjoint kernels cover, for the error of every other plan.
static int adjoint_unsupported() {
  return fail(EXMI_ERR_UNSUPPORTED,
              "no adjoint for this plan (built for ETDRK orders 0-4 in 1-D, 2-D and 3-D with no nonlinearity, "
              "convection, the gradient norm, the polynomial reactions, Cahn-Hilliard, the coupled reactions, "
              "the general form, the 2-D vorticity convection or the projected 3-D convection, real or "
              "complex L)");
}

int exmi_step_vjp(exmi_plan* p, const float* u, int64_t u_bs, const float* g, int64_t g_bs, float* g_in,
                  int64_t gin_bs, void* stream) {
  if (!p || !u || !g || !g_in) return fail(EXMI_ERR_INVALID, "null argument");
  if (!has_adjoint(p))
    return adjoint_unsupported();
  int rc = ensure_adjoint_ws(p);
  if (rc) return rc;
  return enqueue_step_vjp(p, u, u_bs, g, g_bs, g_in, gin_bs, (hipStream_t)stream);
}

int exmi_rollout_vjp(exmi_plan* p, const float* u0, int64_t u0_bs, int32_t n, int32_t include_init,
                     const float* g_traj, const float* g_final, float* g_u0, int32_t checkpoint_every,
                     void* stream) {
  if (!p || !u0 || !g_u0) return fail(EXMI_ERR_INVALID, "null argument");
  if (n < 0) return fail(EXMI_ERR_INVALID, "n_steps must be >= 0");
  if (!has_adjoint(p))
    return adjoint_unsupported();
  hipStream_t st = (hipStream_t)stream;
  const int B = p->d.batch;
  const long long sz = state_size(p);
  const long long bsz = (long long)B * sz;
  const int off = include_init ? 1 : 0;
  const long long gts = (long long)(n + off) * sz;  // g_traj batch stride
  int K = checkpoint_every > 0 && checkpoint_every < n ? checkpoint_every : (n > 0 ? n : 1);
  const int nseg = n > 0 ? (n + K - 1) / K : 0;
  int rc = ensure_adjoint_ws(p);
  if (!rc) rc = ensure_work(p);
  if (rc) return rc;
  // workspace: nseg checkpoints (B, sz) | one segment (B, K, sz) | accumulator (B, sz)
  const size_t need = (size_t)((long long)nseg * bsz + (long long)K * bsz + bsz);
  if (p->ckpt_elems < need) {
    HIP_TRY(hipStreamSynchronize(st));
    hipFree(p->ckpt);
    p->ckpt = nullptr;
    p->ckpt_elems = 0;
    if (hipMalloc(&p->ckpt, need * sizeof(float)) != hipSuccess)
      return fail(EXMI_ERR_ALLOC, "rollout-gradient workspace allocation failed (%zu floats)", need);
    p->ckpt_elems = need;
  }
  // the cooperative reverse pass takes each segment's gradient fields from
  // its forward recomputation (ninst x 4 fields of N^2 per step)
  const bool coop = p->coop_adj && effective_nt(p) == 3 && !g_traj && n > 0;
  const long long gblk = coop ? (long long)p->ninst * 4 * p->d.num_points * p->d.num_points : 0;
  if (coop && p->gseg_elems < (long long)K * gblk) {
    HIP_TRY(hipStreamSynchronize(st));
    hipFree(p->Gseg);
    p->Gseg = nullptr;
    p->gseg_elems = 0;
    if (hipMalloc(&p->Gseg, sizeof(float) * K * gblk) != hipSuccess)
      return fail(EXMI_ERR_ALLOC, "rollout-gradient field workspace allocation failed (%lld floats)", K * gblk);
    p->gseg_elems = K * gblk;
  }
  float* cp = p->ckpt;
  float* seg = cp + (long long)nseg * bsz;
  float* acc = seg + (long long)K * bsz;
  hipError_t e = g_final ? copy_batch(g_final, sz, acc, sz, sz, B, st) : fill(acc, bsz, 0.f, st);
  if (e != hipSuccess) return fail(EXMI_ERR_HIP, "launch failed: %s", hipGetErrorString(e));
  if (n > 0) {
    // forward pass keeping every K-th state
    e = copy_batch(u0, u0_bs, cp, sz, sz, B, st);
    if (e != hipSuccess) return fail(EXMI_ERR_HIP, "launch failed: %s", hipGetErrorString(e));
    for (int sidx = 1; sidx < nseg; ++sidx) {
      rc = enq_repeat_cb(p, cp + (long long)(sidx - 1) * bsz, cp + (long long)sidx * bsz, sz, K, 0, st);
      if (rc) return rc;
    }
    // backward over the segments, recomputing each segment's states
    for (int sidx = nseg - 1; sidx >= 0; --sidx) {
      const int t0 = sidx * K;
      const int len = (n - t0) < K ? (n - t0) : K;
      if (coop) {
        // the segment's len steps as one cooperative launch that keeps the
        // gradient fields of every step (its states are not needed; the last
        // one lands in seg), then the transposed steps from those fields
        rc = enqueue_coop2d(p, cp + (long long)sidx * bsz, sz, seg, sz, 0, len, 1, st, p->Gseg, gblk);
        if (rc) return rc;
      } else {
        rc = enqueue_rollout_nt(p, effective_nt(p), cp + (long long)sidx * bsz, sz, seg, len - 1, 1, st);
        if (rc) return rc;
      }
      const long long segs = (long long)len * sz;  // segment batch stride
      if (coop) {
        // the segment's transposed steps as one cooperative launch (the
        // plan's forward runs cooperatively too)
        CoopAdj2DArgs ca;
        memset(&ca, 0, sizeof ca);
        fill_args2d(p, &ca.aa.s);
        ca.aa.G = p->G;
        ca.aa.LAM = p->LAM;
        ca.aa.RHO = p->RHO;
        ca.aa.H = p->H;
        ca.aa.P = p->P;
        ca.Gseg = p->Gseg;
        ca.G_tstride = gblk;
        ca.seg = seg;
        ca.seg_bstride = segs;
        ca.seg_tstride = sz;
        ca.len = len;
        ca.acc = acc;
        ca.acc_bstride = sz;
        ca.ninst = p->ninst;
        ca.sync = p->coop_sync;
        ca.spin_limit = p->coop_spin;
        e = p->coop_adj(ca, st);
        if (e != hipSuccess) return fail(EXMI_ERR_HIP, "cooperative adjoint launch failed: %s", hipGetErrorString(e));
        continue;
      }
      for (int i = len - 1; i >= 0; --i) {
        const int t = t0 + i;
        if (g_traj) {
          e = axpy_batch(acc, sz, g_traj + (long long)(t + off) * sz, gts, sz, B, st);
          if (e != hipSuccess) return fail(EXMI_ERR_HIP, "launch failed: %s", hipGetErrorString(e));
        }
        rc = enqueue_step_vjp(p, seg + (long long)i * sz, segs, acc, sz, acc, sz, st);
        if (rc) return rc;
      }
    }
  }
  if (g_traj && include_init) {
    e = axpy_batch(acc, sz, g_traj, gts, sz, B, st);
    if (e != hipSuccess) return fail(EXMI_ERR_HIP, "launch failed: %s", hipGetErrorString(e));
  }
  e = copy_batch(acc, sz, g_u0, sz, sz, B, st);
  return e == hipSuccess ? EXMI_OK : fail(EXMI_ERR_HIP, "launch failed: %s", hipGetErrorString(e));
}

}  // extern "C"
