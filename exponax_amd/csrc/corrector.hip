// Kernels and C ABI of the learned corrector (see corrector.hpp): forward
// (inference, or training with stored intermediates), input-gradient pass,
// weight-gradient passes and the deterministic reduction of their per-wave
// partial sums.
#include <hip/hip_runtime.h>

#include "../../include/exmi.h"
#include "corrector.hpp"

namespace exmi {
int set_error(int code, const char* msg);
}

namespace exmi {
namespace cnn {

// Fragment packing: forward A operands W[co][ci][k] and transposed A
// operands W[co][ci][k] with the roles of co and ci exchanged.
__global__ void pack_kernel(const float* __restrict__ p, bf16x8* __restrict__ frag) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NFRAG) return;
  bf16x8 v;
  if (idx >= FRAG_L5) {  // layer 5: A[row 0][K = (tap, ci)] = W5[ci][k], other rows 0
    const int l = (idx - FRAG_L5) % 64, s = (idx - FRAG_L5) / 64, k = s >> 1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int ci = (s & 1) * 16 + 8 * (l >> 5) + i;
      v[i] = static_cast<__bf16>((l & 31) == 0 ? p[OW5 + ci * KW + k] : 0.f);
    }
    frag[idx] = v;
    return;
  }
  if (idx >= FRAG_L1) {  // layer 1: A[co][K = tap] = W1[co][k] for k < 5, else 0
    const int l = idx - FRAG_L1, co = l & 31;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = 8 * (l >> 5) + i;
      v[i] = static_cast<__bf16>(k < KW ? p[OW1 + co * KW + k] : 0.f);
    }
    frag[idx] = v;
    return;
  }
  const int l = idx % 64, s = (idx / 64) % KSTEPS, LT = idx / (64 * KSTEPS);
  const bool T = LT >= 3;
  const int L = T ? LT - 3 : LT;
  const int ow = L == 0 ? OW2 : (L == 1 ? OW3 : OW4);
  const int row = l & 31, k = s >> 1;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int col = (s & 1) * 16 + 8 * (l >> 5) + i;
    const int co = T ? col : row, ci = T ? row : col;
    v[i] = static_cast<__bf16>(p[ow + (co * CH + ci) * KW + k]);
  }
  frag[idx] = v;
}

struct Args {
  const float* params;
  const bf16x8* frag;  // packed fragments (pack_kernel)
  const float* x;      // forward input, sample i at x + i * xs
  float* y;            // forward output, sample i at y + i * ys
  long long xs, ys;    // sample strides in floats (N when contiguous)
  const float* dy;     // backward: cotangent of y (count, N)
  float* dx;           // backward: cotangent of x, or null
  Workspace ws;        // training intermediates, or all null
  float* partial;      // weight-gradient partial sums [waves][NPARAM]
  long long count;
};

template <int N>
struct Lds {
  Edge e;
  bf16x8 frag[3 * KSTEPS * 64];
  __attribute__((aligned(16))) __bf16 act[WAVES][2][N * LDA];
  float vec[WAVES][N];
};

__device__ inline void load_frags(bf16x8* dst, const bf16x8* __restrict__ src, bool transposed) {
  for (int i = threadIdx.x; i < 3 * KSTEPS * 64; i += blockDim.x) dst[i] = src[(transposed ? 3 * KSTEPS * 64 : 0) + i];
}

// The forward pass's workgroup: where one layer's output tiles are all
// formed before any is written (N / 32 tiles, all taken together), a wave
// keeps one activation buffer and overwrites it in place (its reads of a
// layer's input precede its writes of the output in program order, and a
// wave's LDS operations complete in order); the LDS saved holds eight waves
// per workgroup, two workgroups per CU: four waves per SIMD to overlap one
// wave's GELU with another's matrix-core work.
template <int N>
struct FwdCfg {
  static constexpr int NB = N / 32 == tiles_together<N>() ? 1 : 2;  // activation buffers per wave
  static constexpr int NW = NB == 1 ? 8 : WAVES;                    // waves per workgroup
};

template <int N>
struct LdsFwd {
  Edge e;
  bf16x8 frag[3 * KSTEPS * 64];
  __attribute__((aligned(16))) __bf16 act[FwdCfg<N>::NW][FwdCfg<N>::NB][N * LDA];
  float vec[FwdCfg<N>::NW][N];
};

template <int N, bool TRAIN>
__global__ void __launch_bounds__(64 * FwdCfg<N>::NW) forward_kernel(Args a) {
  constexpr int NW = FwdCfg<N>::NW, NB = FwdCfg<N>::NB;
  __shared__ LdsFwd<N> sh;
  load_edge(sh.e, a.params);
  load_frags(sh.frag, a.frag, false);
  __syncthreads();
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const bf16x8* W = sh.frag;
  float* xs = sh.vec[w];
  // every wave runs the same bounded loop over its samples
  for (long long smp = (long long)blockIdx.x * NW + w; smp < a.count; smp += (long long)gridDim.x * NW) {
    for (int p = lane; p < N; p += 64) xs[p] = a.x[smp * a.xs + p];
    wave_sync();
    forward_sample<N, TRAIN>(sh.e, W, a.frag + FRAG_L1, a.frag + FRAG_L5, xs, sh.act[w][0], sh.act[w][NB - 1],
                             a.y + smp * a.ys, a.ws, smp, a.count, lane);
  }
}

// Input-gradient pass: deltas of layers 4..1 (stored for the weight pass)
// and dx = dy + the transposed first layer.
template <int N>
__global__ void __launch_bounds__(64 * WAVES) backward_data_kernel(Args a) {
  __shared__ Lds<N> sh;
  load_edge(sh.e, a.params);
  load_frags(sh.frag, a.frag, true);
  __syncthreads();
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int c = lane & 31, h = lane >> 5;
  const bf16x8* W = sh.frag;
  const Edge& e = sh.e;
  float* dys = sh.vec[w];
  const Workspace ws = a.ws;
  const long long cnt = a.count;
  for (long long smp = (long long)blockIdx.x * WAVES + w; smp < cnt; smp += (long long)gridDim.x * WAVES) {
    for (int p = lane; p < N; p += 64) dys[p] = a.dy[smp * N + p];
    wave_sync();
    __bf16* in = sh.act[w][0];
    __bf16* out = sh.act[w][1];
    // layer 5 transposed (VALU): da4[p][ci] = sum_k W5[ci][k] dy[p - k + 2]
    for (int p = lane; p < N; p += 64) {
      float dv[KW];
#pragma unroll
      for (int k = 0; k < KW; ++k) dv[k] = dys[(p - k + 2) & (N - 1)];
      __bf16 outv[CH];
#pragma unroll
      for (int ci = 0; ci < CH; ++ci) {
        float da = 0.f;
#pragma unroll
        for (int k = 0; k < KW; ++k) da = fmaf(e.w5[ci * KW + k], dv[k], da);
        const long long o = ((3 * cnt + smp) * CH + ci) * N + p;
        const __bf16 d = static_cast<__bf16>(da * gelu_grad(ws.Z[o]));
        outv[ci] = d;
        ws.D[o] = d;
      }
#pragma unroll
      for (int g = 0; g < CH / 8; ++g) {
        bf16x8 v;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = outv[8 * g + i];
        *reinterpret_cast<bf16x8*>(in + p * LDA + 8 * g) = v;
      }
    }
    wave_sync();
    // layers 4, 3, 2 transposed on the matrix cores -> deltas of layers 3, 2, 1
    constexpr int NTL = tiles_together<N>();
#pragma unroll
    for (int L = 2; L >= 0; --L) {
#pragma unroll
      for (int t0 = 0; t0 < N / 32; t0 += NTL) {
        f32x16 acc[NTL];
        conv_tiles<N, true, NTL>(acc, W + L * KSTEPS * 64, in, t0, lane);
#pragma unroll
        for (int q = 0; q < NTL; ++q) {
          const int pos = (t0 + q) * 32 + c;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            bf16x4 v;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int ci = 8 * j + 4 * h + r;
              const long long o = (((long long)L * cnt + smp) * CH + ci) * N + pos;
              const __bf16 d = static_cast<__bf16>(acc[q][4 * j + r] * gelu_grad(ws.Z[o]));
              v[r] = d;
              ws.D[o] = d;
            }
            *reinterpret_cast<bf16x4*>(out + pos * LDA + 8 * j + 4 * h) = v;
          }
        }
      }
      wave_sync();
      __bf16* tmp = in;
      in = out;
      out = tmp;
    }
    // layer 1 transposed (VALU): dx = dy + sum_{co,k} W1[co][k] d1[p - k + 2][co]
    if (a.dx) {
      for (int p = lane; p < N; p += 64) {
        float g = dys[p];
#pragma unroll
        for (int k = 0; k < KW; ++k) {
          const __bf16* row = in + ((p - k + 2) & (N - 1)) * LDA;
#pragma unroll
          for (int q = 0; q < CH / 8; ++q) {
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(row + 8 * q);
#pragma unroll
            for (int i = 0; i < 8; ++i) g = fmaf(e.w1[(8 * q + i) * KW + k], static_cast<float>(v[i]), g);
          }
        }
        a.dx[smp * N + p] = g;
      }
    }
    wave_sync();
  }
}

// Weight gradients of the 32 -> 32 layers on the matrix cores:
// dW[co][ci][k] = sum_{samples, p} d[co][p] a[ci][(p + k - 2) mod N],
// M = co, N = ci (one tile per k), K = positions. Bias gradients on the
// VALU. Each wave writes its partial sums for layers 2-4.
template <int N>
__global__ void __launch_bounds__(64 * WAVES) wgrad_hidden_kernel(Args a) {
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int c = lane & 31, h = lane >> 5;
  const long long wave = (long long)blockIdx.x * WAVES + w, nw = (long long)gridDim.x * WAVES;
  const Workspace ws = a.ws;
  const long long cnt = a.count;
  float* part = a.partial + wave * NPARAM;
  for (int L = 0; L < 3; ++L) {
    f32x16 acc[KW];
#pragma unroll
    for (int k = 0; k < KW; ++k) acc[k] = f32x16{};
    float bsum = 0.f;
    for (long long smp = wave; smp < cnt; smp += nw) {
      const __bf16* D = ws.D + (((long long)(L + 1) * cnt + smp) * CH) * N;  // delta of layer L + 2
      const __bf16* A = ws.A + (((long long)L * cnt + smp) * CH) * N;        // its input a_{L+1}
#pragma unroll
      for (int q = 0; q < N / 16; ++q) {
        const int p0 = 16 * q + 8 * h;
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(D + c * N + p0);
#pragma unroll
        for (int i = 0; i < 8; ++i) bsum += static_cast<float>(av[i]);
#pragma unroll
        for (int k = 0; k < KW; ++k) {
          bf16x8 bv;
#pragma unroll
          for (int i = 0; i < 8; ++i) bv[i] = A[c * N + ((p0 + i + k - 2) & (N - 1))];
          acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[k], 0, 0, 0);
        }
      }
    }
    const int ow = L == 0 ? OW2 : (L == 1 ? OW3 : OW4);
    const int ob = L == 0 ? OB2 : (L == 1 ? OB3 : OB4);
#pragma unroll
    for (int k = 0; k < KW; ++k)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int co = 8 * (i >> 2) + 4 * h + (i & 3), ci = c;
        part[ow + (co * CH + ci) * KW + k] = acc[k][i];
      }
    // bias: lanes c and c + 32 hold the two halves of the positions of channel c
    bsum += __shfl_xor(bsum, 32);
    if (h == 0) part[ob + c] = bsum;
  }
}

// Weight gradients of the 1 -> 32 and 32 -> 1 layers (VALU):
// dW1[co][k] = sum d1[co][p] a0[p + k - 2], db1[co] = sum d1[co][p],
// dW5[ci][k] = sum dy[p] a4[ci][p + k - 2], db5 = sum dy[p].
template <int N>
__global__ void __launch_bounds__(64 * WAVES) wgrad_edge_kernel(Args a) {
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long long wave = (long long)blockIdx.x * WAVES + w, nw = (long long)gridDim.x * WAVES;
  const Workspace ws = a.ws;
  const long long cnt = a.count;
  float* part = a.partial + wave * NPARAM;
  // outputs j = lane + 64 m: 0..159 dW1, 160..191 db1, 192..351 dW5, 352 db5
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (long long smp = wave; smp < cnt; smp += nw) {
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int j = lane + 64 * m;
      float s = 0.f;
      if (j < 160) {
        const int co = j / KW, k = j % KW;
        const __bf16* d = ws.D + (smp * CH + co) * N;
        const __bf16* x0 = ws.A0 + smp * N;
        for (int p = 0; p < N; ++p) s = fmaf(static_cast<float>(d[p]), static_cast<float>(x0[(p + k - 2) & (N - 1)]), s);
      } else if (j < 192) {
        const __bf16* d = ws.D + (smp * CH + (j - 160)) * N;
        for (int p = 0; p < N; ++p) s += static_cast<float>(d[p]);
      } else if (j < 352) {
        const int ci = (j - 192) / KW, k = (j - 192) % KW;
        const __bf16* a4 = ws.A + ((3 * cnt + smp) * CH + ci) * N;
        const float* dy = a.dy + smp * N;
        for (int p = 0; p < N; ++p) s = fmaf(dy[p], static_cast<float>(a4[(p + k - 2) & (N - 1)]), s);
      } else if (j == 352) {
        const float* dy = a.dy + smp * N;
        for (int p = 0; p < N; ++p) s += dy[p];
      }
      acc[m] += s;
    }
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const int j = lane + 64 * m;
    if (j < 160) part[OW1 + j] = acc[m];
    else if (j < 192) part[OB1 + j - 160] = acc[m];
    else if (j < 352) part[OW5 + j - 192] = acc[m];
    else if (j == 352) part[OB5] = acc[m];
  }
}

// Two-stage fixed-order reduction of the per-wave partial sums: stage 1
// sums chunks of RCHUNK waves (grid.y = number of chunks), stage 2 sums the
// chunk results. Deterministic for a given wave count.
constexpr int RCHUNK = 16;

__global__ void reduce_chunks_kernel(const float* __restrict__ partial, long long nw, float* __restrict__ stage) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NPARAM) return;
  const long long w0 = (long long)blockIdx.y * RCHUNK;
  const long long w1 = w0 + RCHUNK < nw ? w0 + RCHUNK : nw;
  float s = 0.f;
  for (long long w = w0; w < w1; ++w) s += partial[w * NPARAM + j];
  stage[(long long)blockIdx.y * NPARAM + j] = s;
}

__global__ void reduce_kernel(const float* __restrict__ stage, long long nchunks, float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NPARAM) return;
  float s = 0.f;
  for (long long c = 0; c < nchunks; ++c) s += stage[c * NPARAM + j];
  out[j] = s;
}

int grid_for(long long count, int waves = WAVES) {
  long long g = (count + waves * 4 - 1) / (waves * 4);  // about 4 samples per wave
  if (g > 2048) g = 2048;
  return g < 1 ? 1 : (int)g;
}

template <int N>
hipError_t launch_forward(const Args& a, bool train, hipStream_t st) {
  constexpr int NW = FwdCfg<N>::NW;
  if (train)
    hipLaunchKernelGGL((forward_kernel<N, true>), dim3(grid_for(a.count, NW)), dim3(64 * NW), 0, st, a);
  else
    hipLaunchKernelGGL((forward_kernel<N, false>), dim3(grid_for(a.count, NW)), dim3(64 * NW), 0, st, a);
  return hipGetLastError();
}

template <int N>
hipError_t launch_backward(const Args& a, int wgrad_groups, hipStream_t st) {
  hipLaunchKernelGGL((backward_data_kernel<N>), dim3(grid_for(a.count)), dim3(64 * WAVES), 0, st, a);
  hipLaunchKernelGGL((wgrad_hidden_kernel<N>), dim3(wgrad_groups), dim3(64 * WAVES), 0, st, a);
  hipLaunchKernelGGL((wgrad_edge_kernel<N>), dim3(wgrad_groups), dim3(64 * WAVES), 0, st, a);
  return hipGetLastError();
}

bool size_ok(int N) { return N == 32 || N == 64 || N == 128; }

}  // namespace cnn

// Internal entry points for the corrected rollout in abi.hip.
size_t cnn_frag_bytes() { return sizeof(cnn::bf16x8) * cnn::NFRAG; }

hipError_t cnn_pack(const float* params, void* frag, hipStream_t st) {
  hipLaunchKernelGGL(cnn::pack_kernel, dim3((cnn::NFRAG + 255) / 256), dim3(256), 0, st, params,
                     static_cast<cnn::bf16x8*>(frag));
  return hipGetLastError();
}

hipError_t cnn_forward_strided(const float* params, const void* frag, int N, long long count, const float* x,
                               long long xs, float* y, long long ys, hipStream_t st) {
  cnn::Args a{};
  a.params = params;
  a.frag = static_cast<const cnn::bf16x8*>(frag);
  a.x = x;
  a.y = y;
  a.xs = xs;
  a.ys = ys;
  a.count = count;
  return N == 32 ? cnn::launch_forward<32>(a, false, st)
                 : (N == 64 ? cnn::launch_forward<64>(a, false, st) : cnn::launch_forward<128>(a, false, st));
}

}  // namespace exmi

using namespace exmi::cnn;

extern "C" {

int64_t exmi_cnn_workspace_bytes(int32_t num_points, int32_t count) {
  return workspace_bytes(num_points, count);
}

int exmi_cnn_forward(const float* params, int32_t N, int32_t count, const float* x, float* y, void* workspace,
                     void* stream) {
  if (!params || !x || !y) return exmi::set_error(EXMI_ERR_INVALID, "exmi_cnn_forward: null argument");
  if (!size_ok(N)) return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_cnn_forward: num_points must be 32, 64 or 128");
  if (count <= 0) return EXMI_OK;
  hipStream_t st = (hipStream_t)stream;
  bf16x8* frag = nullptr;
  if (hipMallocAsync((void**)&frag, sizeof(bf16x8) * NFRAG, st) != hipSuccess)
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_cnn_forward: fragment buffer allocation failed");
  hipLaunchKernelGGL(pack_kernel, dim3((NFRAG + 255) / 256), dim3(256), 0, st, params, frag);
  Args a{};
  a.params = params;
  a.frag = frag;
  a.x = x;
  a.y = y;
  a.xs = a.ys = N;
  a.count = count;
  const bool train = workspace != nullptr;
  if (train) a.ws = workspace_at(workspace, N, count);
  hipError_t e = N == 32 ? launch_forward<32>(a, train, st)
                         : (N == 64 ? launch_forward<64>(a, train, st) : launch_forward<128>(a, train, st));
  (void)hipFreeAsync(frag, st);
  return e == hipSuccess ? EXMI_OK : exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
}

int exmi_cnn_backward(const float* params, int32_t N, int32_t count, const float* dy, float* dx, float* dparams,
                      void* workspace, void* stream) {
  if (!params || !dy || !dparams || !workspace)
    return exmi::set_error(EXMI_ERR_INVALID, "exmi_cnn_backward: null argument");
  if (!size_ok(N)) return exmi::set_error(EXMI_ERR_UNSUPPORTED, "exmi_cnn_backward: num_points must be 32, 64 or 128");
  if (count <= 0) return EXMI_OK;
  hipStream_t st = (hipStream_t)stream;
  const int groups = count < 4 * WAVES * 64 ? (count + WAVES - 1) / WAVES : 64;  // weight-pass workgroups
  const long long nw = (long long)groups * WAVES;
  bf16x8* frag = nullptr;
  float* partial = nullptr;
  if (hipMallocAsync((void**)&frag, sizeof(bf16x8) * NFRAG, st) != hipSuccess)
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_cnn_backward: fragment buffer allocation failed");
  const long long nchunks = (nw + RCHUNK - 1) / RCHUNK;
  if (hipMallocAsync((void**)&partial, sizeof(float) * (nw + nchunks) * NPARAM, st) != hipSuccess) {
    (void)hipFreeAsync(frag, st);
    return exmi::set_error(EXMI_ERR_ALLOC, "exmi_cnn_backward: partial-sum allocation failed");
  }
  hipLaunchKernelGGL(pack_kernel, dim3((NFRAG + 255) / 256), dim3(256), 0, st, params, frag);
  Args a{};
  a.params = params;
  a.frag = frag;
  a.dy = dy;
  a.dx = dx;
  a.ws = workspace_at(workspace, N, count);
  a.partial = partial;
  a.count = count;
  hipError_t e = N == 32 ? launch_backward<32>(a, groups, st)
                         : (N == 64 ? launch_backward<64>(a, groups, st) : launch_backward<128>(a, groups, st));
  if (e == hipSuccess) {
    float* stage = partial + nw * NPARAM;
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3((NPARAM + 255) / 256, (unsigned)nchunks), dim3(256), 0, st, partial,
                       nw, stage);
    hipLaunchKernelGGL(reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, stage, nchunks, dparams);
    e = hipGetLastError();
  }
  (void)hipFreeAsync(frag, st);
  (void)hipFreeAsync(partial, st);
  return e == hipSuccess ? EXMI_OK : exmi::set_error(EXMI_ERR_HIP, hipGetErrorString(e));
}

}  // extern "C"
