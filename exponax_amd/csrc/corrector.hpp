// Learned corrector: a periodic 1-D CNN (5 Conv1d layers, kernel 5,
// 1 -> 32 -> 32 -> 32 -> 32 -> 1 channels, GELU in tanh form after layers
// 1-4, residual output y = x + f(x)) on N points (N in {32, 64, 128}).
//
// Arithmetic: weights and the activations entering each convolution are
// bfloat16; products accumulate in float32; biases, GELU and the residual
// are float32. In the forward pass every layer runs on the gfx950 matrix
// cores (v_mfma_f32_32x32x16_bf16): the three 32 -> 32 layers (98% of the
// flops), the 1 -> 32 layer as one K-step whose K = (tap) holds the 5 taps
// and 11 zeros, and the 32 -> 1 layer as the hidden layers' GEMM with one
// non-zero output row, so the VALU is left with the biases, GELU and the
// residual. The backward passes run the edge layers on the VALU.
//
// One wavefront owns one sample at a time. Its activations live in LDS as
// [position][channel] rows of 40 bfloat16 (32 channels + 8 pad, 80 bytes, so
// each 8-channel group is one aligned 16-byte read). The convolution
// out[co][p] = sum_{ci,k} W[co][ci][k] in[ci][(p + k - 2) mod N] is a GEMM
// with M = co, N = p and K = (k, ci) ordered k-major, so one K-step of 16
// is 16 channels at one shift k:
//   A (32 x 16): lane l holds W[co = l % 32][ci = 16 (s & 1) + 8 (l / 32) + i][k = s / 2], i < 8
//   B (16 x 32): lane l holds in[ci = 16 (s & 1) + 8 (l / 32) + i][p = p0 + l % 32 + k - 2]
//   D (32 x 32): lane l, element i is row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32
// A and B use the same (lane, element) -> K assignment, so the product
// does not depend on the hardware's internal K order.
#pragma once
#include <hip/hip_runtime.h>

namespace exmi {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int CH = 32;
constexpr int KW = 5;
constexpr int LDA = 40;  // bfloat16 per activation row in LDS
// flat parameter vector: W1 (32,1,5) b1 W2 (32,32,5) b2 W3 b3 W4 b4 W5 (1,32,5) b5
constexpr int OW1 = 0, OB1 = 160, OW2 = 192, OB2 = OW2 + 5120, OW3 = OB2 + 32, OB3 = OW3 + 5120;
constexpr int OW4 = OB3 + 32, OB4 = OW4 + 5120, OW5 = OB4 + 32, OB5 = OW5 + 160;
constexpr int NPARAM = OB5 + 1;
static_assert(NPARAM == 15809, "parameter count");
constexpr int KSTEPS = CH * KW / 16;  // 10 MFMA K-steps per 32 -> 32 layer
// packed fragments: 3 forward + 3 transposed hidden layers, then the forward
// edge layers (layer 1: one K-step; layer 5: KSTEPS K-steps, row 0 only)
constexpr int FRAG_L1 = 6 * KSTEPS * 64;
constexpr int FRAG_L5 = FRAG_L1 + 64;
constexpr int NFRAG = FRAG_L5 + KSTEPS * 64;
constexpr int WAVES = 4;                // wavefronts per workgroup

constexpr float GELU_C = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float GELU_A = 0.044715f;

__device__ __forceinline__ float bf16r(float v) { return static_cast<float>(static_cast<__bf16>(v)); }

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = c (z + a z^3), written as
// z * sigmoid(2u) (the same function): sigmoid(2u) = 1 / (1 + 2^(z (K1 + K2 z^2)))
// with K1 = -2 c log2(e), K2 = K1 a, the hardware exp2 and reciprocal
// (7 VALU operations). It saturates to 0 and 1 without overflow trouble:
// rcp(inf) = 0.
constexpr float GELU_K1 = -2.f * GELU_C * 1.4426950408889634f;
constexpr float GELU_K2 = GELU_K1 * GELU_A;
__device__ __forceinline__ float sig2u(float z) {
  const float w = fmaf(GELU_K2, z * z, GELU_K1);
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z * w));
}
__device__ __forceinline__ float gelu(float z) { return z * sig2u(z); }
// d/dz z sigmoid(2u) = s + 2 z s (1 - s) c (1 + 3 a z^2)
__device__ __forceinline__ float gelu_grad(float z) {
  const float s = sig2u(z);
  return fmaf(2.f * z * s * (1.f - s), GELU_C * fmaf(3.f * GELU_A, z * z, 1.f), s);
}

// Edge-layer parameters and hidden biases staged in LDS (weights rounded to
// bfloat16, kept as float).
struct Edge {
  float w1[CH * KW];  // [co][k]
  float b1[CH];
  float w5[CH * KW];  // [ci][k]
  float b5;
  float bh[3][CH];  // b2, b3, b4
};

__device__ inline void load_edge(Edge& e, const float* __restrict__ p) {
  for (int i = threadIdx.x; i < CH * KW; i += blockDim.x) {
    e.w1[i] = bf16r(p[OW1 + i]);
    e.w5[i] = bf16r(p[OW5 + i]);
  }
  for (int i = threadIdx.x; i < CH; i += blockDim.x) {
    e.b1[i] = p[OB1 + i];
    e.bh[0][i] = p[OB2 + i];
    e.bh[1][i] = p[OB3 + i];
    e.bh[2][i] = p[OB4 + i];
  }
  if (threadIdx.x == 0) e.b5 = p[OB5];
}

// Packed fragment index of hidden layer L (0..2 = layers 2..4), K-step s,
// lane l; T selects the transposed (input-gradient) operand.
__host__ __device__ constexpr int frag_index(bool T, int L, int s, int l) {
  return ((T ? 3 + L : L) * KSTEPS + s) * 64 + l;
}

// 32 x 32 output tiles t0 .. t0 + NTL - 1 (positions 32 t ..) of one hidden
// layer from the activation rows `in` (shift +k - 2 forward, -k + 2
// transposed). W: this layer's A fragments in LDS, [s][lane]. The tiles'
// accumulation chains are interleaved to hide the MFMA dependency latency.
template <int N, bool TRANSPOSED, int NTL>
__device__ __forceinline__ void conv_tiles(f32x16 (&acc)[NTL], const bf16x8* W, const __bf16* in, int t0,
                                           int lane) {
  const int c = lane & 31, h = lane >> 5;
#pragma unroll
  for (int q = 0; q < NTL; ++q) acc[q] = f32x16{};
#pragma unroll
  for (int s = 0; s < KSTEPS; ++s) {
    const int k = s >> 1;
    const bf16x8 a = W[s * 64 + lane];
#pragma unroll
    for (int q = 0; q < NTL; ++q) {
      const int p = ((t0 + q) * 32 + c + (TRANSPOSED ? 2 - k : k - 2)) & (N - 1);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(in + p * LDA + (s & 1) * 16 + 8 * h);
      acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
    }
  }
}

// tiles processed together: two when the layer has an even number of tiles
template <int N>
constexpr int tiles_together() {
  return (N / 32) % 2 == 0 ? 2 : 1;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Training workspace of `count` samples (see exmi_cnn_workspace_bytes):
//   Z  float [4][count][32][N]  pre-activations z1..z4
//   A0 bf16  [count][N]         rounded input
//   A  bf16  [4][count][32][N]  activations a1..a4
//   D  bf16  [4][count][32][N]  deltas (gradients of z1..z4)
struct Workspace {
  float* Z;
  __bf16* A0;
  __bf16* A;
  __bf16* D;
};

__host__ __device__ inline long long workspace_bytes(int N, long long count) {
  return count * N * (4LL * CH * 4 + 2 + 4LL * CH * 2 + 4LL * CH * 2);
}

__host__ __device__ inline Workspace workspace_at(void* base, int N, long long count) {
  char* b = static_cast<char*>(base);
  Workspace w;
  w.Z = reinterpret_cast<float*>(b);
  b += 4LL * count * CH * N * 4;
  w.A0 = reinterpret_cast<__bf16*>(b);
  b += count * N * 2;
  w.A = reinterpret_cast<__bf16*>(b);
  b += 4LL * count * CH * N * 2;
  w.D = reinterpret_cast<__bf16*>(b);
  return w;
}

// Forward of one sample by one wavefront. xs: the sample's N inputs in LDS
// (float); act0/act1: this wave's two activation buffers; the result
// y = x + f(x) is written to y (global). With `ws` the training quantities
// of sample `smp` are stored.
// Bias, GELU and the bfloat16 activation rows of 32 x 32 output tiles
// t0.. (layer index Lz of the training workspace).
template <int N, bool TRAIN, int NTL>
__device__ __forceinline__ void act_epilogue(const f32x16 (&acc)[NTL], const float* bias, __bf16* out, int t0,
                                             int lane, Workspace ws, long long smp, long long count, int Lz) {
  const int c = lane & 31, h = lane >> 5;
#pragma unroll
  for (int q = 0; q < NTL; ++q) {
    const int pos = (t0 + q) * 32 + c;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bf16x4 v;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = 8 * j + 4 * h + r;
        const float z = acc[q][4 * j + r] + bias[co];
        v[r] = static_cast<__bf16>(gelu(z));
        if constexpr (TRAIN) {
          const long long o = (((long long)Lz * count + smp) * CH + co) * N + pos;
          ws.Z[o] = z;
          ws.A[o] = v[r];
        }
      }
      *reinterpret_cast<bf16x4*>(out + pos * LDA + 8 * j + 4 * h) = v;
    }
  }
}

// Forward of one sample by one wavefront. xs: the sample's N inputs in LDS
// (float); act0/act1: this wave's two activation buffers (one and the same
// when every output tile of a layer is formed before any is written); W: the hidden
// layers' A fragments (LDS); F1, F5: the edge layers' A fragments (global,
// cached); the result y = x + f(x) is written to y (global). With `ws` the
// training quantities of sample `smp` are stored.
template <int N, bool TRAIN>
__device__ __forceinline__ void forward_sample(const Edge& e, const bf16x8* W, const bf16x8* F1, const bf16x8* F5,
                                               const float* xs, __bf16* act0, __bf16* act1, float* __restrict__ y,
                                               Workspace ws, long long smp, long long count, int lane) {
  const int c = lane & 31, h = lane >> 5;
  constexpr int NTL = tiles_together<N>();
  if constexpr (TRAIN) {
    for (int p = lane; p < N; p += 64) ws.A0[smp * N + p] = static_cast<__bf16>(xs[p]);
  }
  // layer 1 (1 -> 32): K = tap 0..4 (lanes of the upper half and taps >= 5 hold zeros)
  {
    const bf16x8 a1 = F1[lane];
#pragma unroll
    for (int t0 = 0; t0 < N / 32; t0 += NTL) {
      f32x16 acc[NTL];
#pragma unroll
      for (int q = 0; q < NTL; ++q) {
        const int pos = (t0 + q) * 32 + c;
        bf16x8 b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = 8 * h + i;
          b[i] = k < KW ? static_cast<__bf16>(xs[(pos + k - 2) & (N - 1)]) : static_cast<__bf16>(0.f);
        }
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, f32x16{}, 0, 0, 0);
      }
      act_epilogue<N, TRAIN, NTL>(acc, e.b1, act0, t0, lane, ws, smp, count, 0);
    }
  }
  wave_sync();
  // layers 2-4 (32 -> 32)
  __bf16* in = act0;
  __bf16* out = act1;
#pragma unroll
  for (int L = 0; L < 3; ++L) {
#pragma unroll
    for (int t0 = 0; t0 < N / 32; t0 += NTL) {
      f32x16 acc[NTL];
      conv_tiles<N, false, NTL>(acc, W + L * KSTEPS * 64, in, t0, lane);
      act_epilogue<N, TRAIN, NTL>(acc, e.bh[L], out, t0, lane, ws, smp, count, L + 1);
    }
    wave_sync();
    __bf16* tmp = in;
    in = out;
    out = tmp;
  }
  // layer 5 (32 -> 1): output row 0 of the hidden layers' GEMM, residual output
#pragma unroll
  for (int t0 = 0; t0 < N / 32; t0 += NTL) {
    f32x16 acc[NTL];
    conv_tiles<N, false, NTL>(acc, F5, in, t0, lane);
    if (h == 0) {
#pragma unroll
      for (int q = 0; q < NTL; ++q) {
        const int pos = (t0 + q) * 32 + c;
        y[pos] = xs[pos] + (e.b5 + acc[q][0]);
      }
    }
  }
  wave_sync();
}

}  // namespace cnn
}  // namespace exmi
