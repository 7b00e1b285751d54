// GPU counterparts of the functions on either side of the ETDRK path that
// measure states or change their resolution:
//   nRMSE                    metrics/_spatial.py:537-586 (spatial_norm, mode "normalized")
//   first non-finite / out-of-bound state of every trajectory (failure
//   detection; no reference counterpart)
//   get_spectrum (1-D)       _spectral.py:866-1030
//   map_between_resolutions  _interpolation.py:195-285 (1-D)
// Transforms go through the library's own real FFT kernels (exmi_rfft /
// exmi_irfft); the kernels here are the reductions and per-mode maps
// between them. Workspaces are stream-ordered allocations.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/exmi.h"

namespace exmi {
int set_error(int code, const char* msg);  // abi.hip: the thread-local error channel
}

namespace {

constexpr int WG = 256;

// out[b] = sum_c ||p_bc - r_bc|| / ||r_bc|| with float64 accumulation.
__global__ void __launch_bounds__(WG) nrmse_kernel(const float* __restrict__ p, long long pbs,
                                                   const float* __restrict__ r, long long rbs, int C, long long n,
                                                   float* __restrict__ out) {
  __shared__ double sh[2][WG];
  const int b = blockIdx.x;
  double total = 0.0;
  for (int c = 0; c < C; ++c) {
    const float* pc = p + (long long)b * pbs + (long long)c * n;
    const float* rc = r + (long long)b * rbs + (long long)c * n;
    double se = 0.0, sr = 0.0;
    for (long long i = threadIdx.x; i < n; i += WG) {
      const double d = (double)pc[i] - (double)rc[i];
      se += d * d;
      sr += (double)rc[i] * (double)rc[i];
    }
    sh[0][threadIdx.x] = se;
    sh[1][threadIdx.x] = sr;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
        sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
      }
      __syncthreads();
    }
    total += sqrt(sh[0][0]) / sqrt(sh[1][0]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = (float)total;
}

// 1-D spectrum of X (count, N/2+1) complex: power 0.5 |X|^2 / (s_rec s_norm),
// else |X| / s_rec; s_rec = N at k = 0 and (even N) k = N/2, N/2 elsewhere;
// s_norm = N (build_scaling_array "reconstruction" / "norm_compensation").
__global__ void spectrum_kernel(const float2* __restrict__ X, int n, long long total, int power, float* out) {
  const int nk = n / 2 + 1;
  const float N = (float)n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(i % nk);
    const float2 x = X[i];
    const float mag = sqrtf(x.x * x.x + x.y * x.y);
    const float rec = (k == 0 || 2 * k == n) ? N : 0.5f * N;
    const float m = mag / rec;
    out[i] = power ? 0.5f * m * (mag / N) : m;
  }
}

// Radially binned (power) spectrum of D = 2, 3 fields (_spectral.py:866-1030):
// the per-mode quantity |X| / reconstruction scaling (power: times
// 0.5 |X| / N^D), summed (or averaged) over the modes whose wavenumber norm
// lies in [b - 1/2, b + 1/2), b = 0..N/2; NaN quantities are skipped and an
// empty average is NaN (nansum / nanmean). One workgroup per (field,
// bucket), reduced in a fixed order.
__global__ void __launch_bounds__(WG) radial_spectrum_kernel(const float2* __restrict__ X, int D, int N, int power,
                                                             int average, float* __restrict__ out) {
  const int NZ = N / 2 + 1, nb = N / 2 + 1;
  const long long field = blockIdx.x / nb;
  const int b = (int)(blockIdx.x % nb);
  long long nr = NZ;
  for (int d = 1; d < D; ++d) nr *= N;
  const float2* x = X + field * nr;
  float nd = 1.f;
  for (int d = 0; d < D; ++d) nd *= (float)N;
  __shared__ float ssum[WG];
  __shared__ int scnt[WG];
  float acc = 0.f;
  int cnt = 0;
  for (long long m = threadIdx.x; m < nr; m += WG) {
    const int kz = (int)(m % NZ);
    long long r = m / NZ;
    float k2 = (float)kz * (float)kz;
    float rec = (kz == 0 || 2 * kz == N) ? (float)N : 0.5f * (float)N;
    for (int d = 1; d < D; ++d) {
      const int i = (int)(r % N);
      r /= N;
      const int k = i <= (N - 1) / 2 ? i : i - N;
      k2 += (float)k * (float)k;
      rec *= (float)N;
    }
    const float kn = sqrtf(k2);
    if (kn >= (float)b - 0.5f && kn < (float)b + 0.5f) {
      const float2 v = x[m];
      const float mag = sqrtf(v.x * v.x + v.y * v.y);
      const float q = power ? 0.5f * (mag / rec) * (mag / nd) : mag / rec;
      if (q == q) {
        acc += q;
        ++cnt;
      }
    }
  }
  ssum[threadIdx.x] = acc;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int w = WG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
      scnt[threadIdx.x] += scnt[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[field * nb + b] = average ? (scnt[0] ? ssum[0] / (float)scnt[0] : __int_as_float(0x7fc00000)) : ssum[0];
}

// Spectral resampling of (count, Mo+1) -> (count, Mn+1): copy the modes
// k <= min(No, Nn)/2 times Nn/No (norm compensation), zero the old Nyquist
// before upsampling and the new Nyquist after downsampling (oddball rules).
__global__ void resample_kernel(const float2* __restrict__ A, int Mo, float2* __restrict__ Bn, int Mn, int count,
                                int oddball) {
  const long long total = (long long)count * (Mn + 1);
  const int kmax = Mo < Mn ? Mo : Mn;
  const float f = (float)Mn / (float)Mo;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long t = i / (Mn + 1);
    const int k = (int)(i % (Mn + 1));
    float2 v = make_float2(0.f, 0.f);
    if (k <= kmax) {
      v = A[t * (Mo + 1) + k];
      v.x *= f;
      v.y *= f;
      if (oddball && Mn > Mo && k == Mo) v = make_float2(0.f, 0.f);
      if (oddball && Mo > Mn && k == Mn) v = make_float2(0.f, 0.f);
    }
    Bn[i] = v;
  }
}

// The same for D = 2, 3 (_interpolation.py:195-285 with the mode blocks of
// _spectral.py:530-611): a new mode (kx, [ky,] kz) takes the old mode of the
// same signed wavenumbers when every leading-axis k lies in [-h, h - 1] and
// kz <= h, h = min(No, Nn)/2, times (Nn/No)^D; oddball rules zero the old
// grid's Nyquist modes (any axis) before upsampling and the new grid's after
// downsampling.
__global__ void resample_nd_kernel(const float2* __restrict__ A, int No, float2* __restrict__ Bn, int Nn, int D,
                                   long long count, int oddball) {
  const int NZo = No / 2 + 1, NZn = Nn / 2 + 1;
  long long nrn = NZn, nro = NZo;
  for (int d = 1; d < D; ++d) {
    nrn *= Nn;
    nro *= No;
  }
  const int h = (No < Nn ? No : Nn) / 2;
  float f = 1.f;
  for (int d = 0; d < D; ++d) f *= (float)Nn / (float)No;
  const long long total = count * nrn;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long t = i / nrn;
    long long r = i % nrn;
    const int kz = (int)(r % NZn);
    r /= NZn;
    bool copy = kz <= h;
    bool zero_old = kz == No / 2, zero_new = kz == Nn / 2;
    long long oidx = kz, ostride = NZo;
    for (int d = 1; d < D; ++d) {  // leading axes, innermost first
      const int ii = (int)(r % Nn);
      r /= Nn;
      const int k = ii < Nn / 2 ? ii : ii - Nn;
      copy = copy && k >= -h && k <= h - 1;
      zero_old = zero_old || k == -No / 2;
      zero_new = zero_new || k == -Nn / 2;
      oidx += (long long)(k >= 0 ? k : No + k) * ostride;
      ostride *= No;
    }
    float2 v = make_float2(0.f, 0.f);
    if (copy && !(oddball && Nn > No && zero_old) && !(oddball && No > Nn && zero_new)) {
      v = A[t * nro + oidx];
      v.x *= f;
      v.y *= f;
    }
    Bn[i] = v;
  }
}

// first_bad[b] = the smallest t < T whose state x[b, t, 0:n] has a value
// that is not finite or exceeds `bound` in magnitude, -1 when there is none.
// One workgroup per trajectory; every thread keeps the smallest bad t of the
// elements it reads, then a min-reduction.
__global__ void __launch_bounds__(WG) first_bad_kernel(const float* __restrict__ x, long long bs, long long ts,
                                                       int T, long long n, float bound, int offset,
                                                       int* __restrict__ first_bad) {
  __shared__ int sh[WG];
  const int b = blockIdx.x;
  int best = T;
  for (int t = 0; t < T; ++t) {
    const float* xs = x + (long long)b * bs + (long long)t * ts;
    bool bad = false;
    for (long long i = threadIdx.x; i < n; i += WG) bad |= !(fabsf(xs[i]) <= bound);
    if (bad) {  // this thread's first failing time; the minimum over the threads is the trajectory's
      best = t;
      break;
    }
  }
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int w = WG / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w && sh[threadIdx.x + w] < sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) first_bad[b] = sh[0] < T ? offset + sh[0] : -1;
}

int grid_for(long long n) {
  long long g = (n + WG - 1) / WG;
  if (g > 16384) g = 16384;
  return g < 1 ? 1 : (int)g;
}

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

}  // namespace

namespace exmi {
int launch_first_bad(int batch, int n_times, long long n, const float* x, long long bs, long long ts, float bound,
                     int offset, int* first_bad, hipStream_t st) {
  if (batch == 0) return EXMI_OK;
  hipLaunchKernelGGL(first_bad_kernel, dim3(batch), dim3(WG), 0, st, x, bs, ts, n_times, n, bound, offset, first_bad);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? EXMI_OK : set_error(EXMI_ERR_HIP, hipGetErrorString(e));
}
}  // namespace exmi

extern "C" {

int exmi_nrmse(int32_t batch, int32_t channels, int64_t n, const float* pred, int64_t pred_bstride, const float* ref,
               int64_t ref_bstride, float* out, void* stream) {
  if (!pred || !ref || !out) return exmi::set_error(EXMI_ERR_INVALID, "exmi_nrmse: null argument");
  if (batch < 0 || channels < 1 || n < 1) return exmi::set_error(EXMI_ERR_INVALID, "exmi_nrmse: bad sizes");
  if (batch == 0) return EXMI_OK;
  hipLaunchKernelGGL(nrmse_kernel, dim3(batch), dim3(WG), 0, (hipStream_t)stream, pred, (long long)pred_bstride,
                     ref, (long long)ref_bstride, channels, (long long)n, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? EXMI_OK : exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
}

int exmi_first_bad_step(int32_t batch, int32_t n_times, int64_t n, const float* x, int64_t bstride, int64_t tstride,
                        float bound, int32_t* first_bad, void* stream) {
  if (!x || !first_bad) return exmi::set_error(EXMI_ERR_INVALID, "exmi_first_bad_step: null argument");
  if (batch < 0 || n_times < 0 || n < 1) return exmi::set_error(EXMI_ERR_INVALID, "exmi_first_bad_step: bad sizes");
  if (batch == 0) return EXMI_OK;
  const float b = bound > 0.f ? bound : 3.402823466e38f;  // non-finite values fail every bound
  return exmi::launch_first_bad(batch, n_times, (long long)n, x, (long long)bstride, (long long)tstride, b, 0,
                                (int*)first_bad, (hipStream_t)stream);
}

int exmi_spectrum_binned(int32_t D, int32_t N, int32_t count, const float* x, float* out, int32_t power,
                         int32_t average, void* stream) {
  if (D == 1) return exmi_spectrum(D, N, count, x, out, power, stream);
  if (D != 2 && D != 3) return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_spectrum: num_spatial_dims 1, 2, 3");
  if (!x || !out) return exmi::set_error(EXMI_ERR_INVALID, "exmi_spectrum: null argument");
  if (count <= 0) return EXMI_OK;
  hipStream_t st = (hipStream_t)stream;
  long long nr = N / 2 + 1;
  for (int d = 1; d < D; ++d) nr *= N;
  float2* X = nullptr;
  if (hipMallocAsync((void**)&X, sizeof(float2) * nr * count, st) != hipSuccess)
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_spectrum: workspace allocation failed");
  int rc = exmi_rfft(D, N, count, x, reinterpret_cast<float*>(X), stream);
  if (rc == EXMI_OK) {
    hipLaunchKernelGGL(radial_spectrum_kernel, dim3((unsigned)((long long)count * (N / 2 + 1))), dim3(WG), 0, st, X,
                       D, N, power, average, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
  }
  (void)hipFreeAsync(X, st);
  return rc;
}

int exmi_spectrum(int32_t D, int32_t N, int32_t count, const float* x, float* out, int32_t power, void* stream) {
  if (D == 2 || D == 3) return exmi_spectrum_binned(D, N, count, x, out, power, 0, stream);
  if (D != 1) return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_spectrum: num_spatial_dims 1, 2, 3");
  if (!x || !out) return exmi::set_error(EXMI_ERR_INVALID, "exmi_spectrum: null argument");
  if (count <= 0) return EXMI_OK;
  hipStream_t st = (hipStream_t)stream;
  const int M = N / 2;
  const long long total = (long long)count * (M + 1);
  float2* X = nullptr;
  if (hipMallocAsync((void**)&X, sizeof(float2) * total, st) != hipSuccess)
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_spectrum: workspace allocation failed");
  int rc = exmi_rfft(1, N, count, x, reinterpret_cast<float*>(X), stream);
  if (rc == EXMI_OK) {
    hipLaunchKernelGGL(spectrum_kernel, dim3(grid_for(total)), dim3(WG), 0, st, X, N, total, power, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
  }
  (void)hipFreeAsync(X, st);
  return rc;
}

int exmi_map_resolution(int32_t D, int32_t n_old, int32_t n_new, int32_t count, const float* x, float* y,
                        int32_t oddball_zero, void* stream) {
  if (D == 2 || D == 3) {
    const int lo = D == 2 ? 32 : 16, hi = D == 2 ? 512 : 256;
    if (!x || !y) return exmi::set_error(EXMI_ERR_INVALID, "exmi_map_resolution: null argument");
    if (!pow2(n_old) || !pow2(n_new) || n_old < lo || n_new < lo || n_old > hi || n_new > hi)
      return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_map_resolution: sizes must be powers of two in the "
                                                   "transform range of this dimension");
    if (count <= 0) return EXMI_OK;
    hipStream_t st = (hipStream_t)stream;
    long long no = 1, nn = 1, nro = n_old / 2 + 1, nrn = n_new / 2 + 1;
    for (int d = 0; d < D; ++d) {
      no *= n_old;
      nn *= n_new;
    }
    for (int d = 1; d < D; ++d) {
      nro *= n_old;
      nrn *= n_new;
    }
    if (n_old == n_new) {
      const hipError_t e = hipMemcpyAsync(y, x, sizeof(float) * (size_t)count * no, hipMemcpyDeviceToDevice, st);
      return e == hipSuccess ? EXMI_OK : exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
    }
    float2 *A = nullptr, *Bn = nullptr;
    if (hipMallocAsync((void**)&A, sizeof(float2) * (size_t)count * nro, st) != hipSuccess)
      return exmi::set_error(EXMI_ERR_ALLOC, "exmi_map_resolution: workspace allocation failed");
    if (hipMallocAsync((void**)&Bn, sizeof(float2) * (size_t)count * nrn, st) != hipSuccess) {
      (void)hipFreeAsync(A, st);
      return exmi::set_error(EXMI_ERR_ALLOC, "exmi_map_resolution: workspace allocation failed");
    }
    int rc = exmi_rfft(D, n_old, count, x, reinterpret_cast<float*>(A), stream);
    if (rc == EXMI_OK) {
      hipLaunchKernelGGL(resample_nd_kernel, dim3(grid_for((long long)count * nrn)), dim3(WG), 0, st, A, n_old, Bn,
                         n_new, D, (long long)count, oddball_zero);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
    }
    if (rc == EXMI_OK) rc = exmi_irfft(D, n_new, count, reinterpret_cast<const float*>(Bn), y, stream);
    (void)hipFreeAsync(A, st);
    (void)hipFreeAsync(Bn, st);
    return rc;
  }
  if (D != 1) return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_map_resolution: num_spatial_dims 1, 2, 3");
  if (!x || !y) return exmi::set_error(EXMI_ERR_INVALID, "exmi_map_resolution: null argument");
  if (!pow2(n_old) || !pow2(n_new) || n_old < 16 || n_new < 16 || n_old > 4096 || n_new > 4096)
    return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_map_resolution: sizes must be powers of two in 16..4096");
  if (count <= 0) return EXMI_OK;
  hipStream_t st = (hipStream_t)stream;
  if (n_old == n_new) {
    const hipError_t e = hipMemcpyAsync(y, x, sizeof(float) * (size_t)count * n_old, hipMemcpyDeviceToDevice, st);
    return e == hipSuccess ? EXMI_OK : exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
  }
  const int Mo = n_old / 2, Mn = n_new / 2;
  float2 *A = nullptr, *Bn = nullptr;
  if (hipMallocAsync((void**)&A, sizeof(float2) * (size_t)count * (Mo + 1), st) != hipSuccess)
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_map_resolution: workspace allocation failed");
  if (hipMallocAsync((void**)&Bn, sizeof(float2) * (size_t)count * (Mn + 1), st) != hipSuccess) {
    (void)hipFreeAsync(A, st);
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_map_resolution: workspace allocation failed");
  }
  int rc = exmi_rfft(1, n_old, count, x, reinterpret_cast<float*>(A), stream);
  if (rc == EXMI_OK) {
    const long long total = (long long)count * (Mn + 1);
    hipLaunchKernelGGL(resample_kernel, dim3(grid_for(total)), dim3(WG), 0, st, A, Mo, Bn, Mn, count,
                       oddball_zero ? 1 : 0);
    const hipError_t e = hipGetLastError();
    rc = e == hipSuccess ? EXMI_OK : exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
  }
  if (rc == EXMI_OK) rc = exmi_irfft(1, n_new, count, reinterpret_cast<const float*>(Bn), y, stream);
  (void)hipFreeAsync(A, st);
  (void)hipFreeAsync(Bn, st);
  return rc;
}

}  // extern "C"
