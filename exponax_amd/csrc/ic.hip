// Seeded truncated-Fourier-series initial conditions on the device
// (ic/_truncated_fourier_series.py:17-100 semantics with the library's own
// seeded draws). The host draws one amplitude a and phase phi per retained
// mode and field (the same stream as the host recipe); the device writes the
// spectrum of sum_m a_m cos(2 pi k_m . x / L + phi_m), inverse-transforms it
// (exmi_irfft) and divides every field by its max |u|.
#include <hip/hip_runtime.h>

#include "../../include/exmi.h"
#include "exmi_internal.hpp"

namespace exmi {
int set_error(int code, const char* msg);

namespace {

constexpr int WG = 256;

__global__ void zero_kernel(float2* X, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    X[i] = make_float2(0.f, 0.f);
}

// Mode m of field f: X[k] = N^D a e^{i phi} / 2; a mode with k_last = 0
// also sets its partner X[-k] = conj(X[k]) (the c2r of the last axis only
// mirrors k_last > 0). Modes are distinct half-plane representatives, so no
// two writes meet.
__global__ void tfs_spectrum_kernel(float2* X, int D, int N, long long nr, int nmodes, const int* __restrict__ modes,
                                    const float* __restrict__ amp, const float* __restrict__ phase, long long nf,
                                    float half_nd) {
  const int NZ = N / 2 + 1;
  const long long total = nf * nmodes;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long f = i / nmodes;
    const int m = (int)(i % nmodes);
    const int* k = modes + (long long)m * D;
    const float a = amp[i] * half_nd, ph = phase[i];
    float s, c;
    sincosf(ph, &s, &c);
    long long pos = k[D - 1], neg = 0;  // linear indices of k and -k ([k_0]...[k_last])
    const bool pair = k[D - 1] == 0;
    long long stride = NZ;
    for (int d = D - 2; d >= 0; --d) {
      const int kd = k[d];
      pos += (long long)(kd >= 0 ? kd : N + kd) * stride;
      neg += (long long)(kd > 0 ? N - kd : -kd) * stride;
      stride *= N;
    }
    X[f * nr + pos] = make_float2(a * c, a * s);
    if (pair) X[f * nr + neg] = make_float2(a * c, -a * s);
  }
}

__global__ void __launch_bounds__(WG) normalise_kernel(float* u, long long n) {
  __shared__ float sh[WG];
  float* x = u + (long long)blockIdx.x * n;
  float mx = 0.f;
  for (long long i = threadIdx.x; i < n; i += WG) mx = fmaxf(mx, fabsf(x[i]));
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int w = WG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  const float r = 1.f / sh[0];
  for (long long i = threadIdx.x; i < n; i += WG) x[i] *= r;
}

unsigned grid_of(long long n) {
  long long g = (n + WG - 1) / WG;
  return (unsigned)(g > 65536 ? 65536 : g);
}

}  // namespace
}  // namespace exmi

extern "C" int exmi_ic_tfs(int32_t D, int32_t N, int32_t count, int32_t nmodes, const int32_t* modes, const float* amp,
                           const float* phase, float* out, void* stream) {
  using namespace exmi;
  if (D < 1 || D > 3) return set_error(EXMI_ERR_UNSUPPORTED, "exmi_ic_tfs: num_spatial_dims 1, 2, 3");
  if (!modes || !amp || !phase || !out) return set_error(EXMI_ERR_INVALID, "exmi_ic_tfs: null argument");
  if (count < 0 || nmodes < 0) return set_error(EXMI_ERR_INVALID, "exmi_ic_tfs: negative counts");
  if (count == 0 || nmodes == 0) return EXMI_OK;
  // the modes (host array) must be distinct half-plane representatives
  // strictly inside the grid's band
  for (long long i = 0; i < (long long)nmodes * D; ++i)
    if (modes[i] <= -N / 2 || modes[i] >= N / 2)
      return set_error(EXMI_ERR_INVALID, "exmi_ic_tfs: every |k_d| must be below num_points / 2");
  for (int m = 0; m < nmodes; ++m) {
    int last = 0;
    for (int d = D - 1; d >= 0 && last == 0; --d) last = modes[(long long)m * D + d];
    if (last <= 0) return set_error(EXMI_ERR_INVALID, "exmi_ic_tfs: modes must be half-plane representatives");
  }
  hipStream_t st = (hipStream_t)stream;
  long long nr = N / 2 + 1, np = N;
  for (int d = 1; d < D; ++d) {
    nr *= N;
    np *= N;
  }
  float2* X = nullptr;
  int* dmodes = nullptr;
  if (hipMallocAsync((void**)&X, sizeof(float2) * nr * count, st) != hipSuccess)
    return set_error(EXMI_ERR_ALLOC, "exmi_ic_tfs: workspace allocation failed");
  if (hipMallocAsync((void**)&dmodes, sizeof(int) * nmodes * D, st) != hipSuccess ||
      hipMemcpyAsync(dmodes, modes, sizeof(int) * nmodes * D, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    (void)hipFreeAsync(X, st);
    if (dmodes) (void)hipFreeAsync(dmodes, st);
    return set_error(EXMI_ERR_ALLOC, "exmi_ic_tfs: mode upload failed");
  }
  hipLaunchKernelGGL(zero_kernel, dim3(grid_of(nr * count)), dim3(WG), 0, st, X, nr * count);
  hipLaunchKernelGGL(tfs_spectrum_kernel, dim3(grid_of((long long)count * nmodes)), dim3(WG), 0, st, X, D, N, nr,
                     nmodes, dmodes, amp, phase, (long long)count, 0.5f * (float)np);
  hipError_t e = hipGetLastError();
  int rc = e == hipSuccess ? exmi_irfft(D, N, count, reinterpret_cast<const float*>(X), out, stream)
                           : set_error(EXMI_ERR_HIP, hipGetErrorString(e));
  if (rc == EXMI_OK) {
    hipLaunchKernelGGL(normalise_kernel, dim3((unsigned)count), dim3(WG), 0, st, out, np);
    e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(EXMI_ERR_HIP, hipGetErrorString(e));
  }
  (void)hipFreeAsync(X, st);
  (void)hipFreeAsync(dmodes, st);
  return rc;
}
