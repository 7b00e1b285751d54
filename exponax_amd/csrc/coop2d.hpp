// Cooperative 2-D ETDRK2 rollout (gfx950): the spectrum-carrying launch
// chain of step2d.hpp run as ONE launch for all steps of a rollout.
//
// The chain's passes (RA CA once, then per step RB|RCB, CB, RB, CCA, and a
// final RC; etdrk/_etdrk_2.py:91-102 with the nonlinear evaluation of
// nonlin_fun/_base.py:99-137) are the same row and column bodies; between
// two passes the workgroups of one instance meet at a barrier instead of a
// kernel boundary. Instance i owns workgroups i CG .. i CG + CG - 1, one per
// column group (CG = column groups per instance); the first RG of them also
// run the row passes (RG = N / rows-per-workgroup <= CG). A workgroup keeps
// the same columns in every column pass, so the spectral state S0, the stage
// state S1 and the first nonlinear term N0 are read back only by the
// workgroup that wrote them; the row spectra / products T and the fields F
// cross workgroups and are accessed coherently (gld/gst with COH).
//
// Barrier: every lane waits for its own stores, the workgroup meets at
// s_barrier, one lane adds 1 to the instance's arrival counter (a relaxed
// agent-scope atomic) and spins until the counter reaches the barrier's
// count of arrivals. As T and F are written and read with coherent accesses,
// the barrier needs no device-scope fence (a release at agent scope would
// write back the whole L2 of the workgroup's XCD at every barrier). The
// spin is bounded: past `spin_limit` polls, or once any workgroup of the
// launch has given up, the workgroup sets the launch's failure word, writes
// NaN to its rows of every state it has not emitted yet and returns, so every
// wave of the grid exits and no stale data passes for a result.
// All workgroups of the grid must be resident at once, which the launcher
// checks against the kernel's occupancy (an instance's workgroups are
// spread over the XCDs by the dispatcher's round robin; the counters and
// the exchanged data live at the device's point of coherence).
//
// Served plans: those of the chain (real L, ETDRK2, dealiased Nyquist row and
// column) with one channel, gradient norm (KS 2-D) or vorticity
// (Kolmogorov), N = 128 and 256, where the row and column passes both run
// 256-thread workgroups.
#pragma once
#include "adjoint2d.hpp"
#include "step2d.hpp"

namespace exmi {

struct Coop2DArgs {
  Step2DArgs s;            // the chain's arguments (tables, twiddles, buffers, scalars)
  const float* u0;         // (ninst, N, N) initial states, stride u0_bstride floats
  long long u0_bstride;
  float* traj;             // output base
  long long traj_bstride;  // floats between instances
  long long traj_tstride;  // floats between consecutive outputs of one instance
  int nsteps;
  int final_only;          // 1: write only the last state, at traj
  int ninst;
  unsigned* sync;          // arrival counter of instance i at sync[64 i] (zero at launch);
                           // sync[64 ninst]: failure word (nonzero: a barrier gave up)
  int spin_limit;          // polls of a barrier before it gives up
  float* G;                // gradient norm: step t's gradient fields of both stages to G + t G_tstride
  long long G_tstride;     // (the adjoint's layout, see Step2DArgs::Gout), or G null
};

template <int N>
struct CoopCfg {
  using R = RowCfg<N>;
  using C = ColCfg<N>;
  static constexpr int CG = (C::NY + C::CPB - 1) / C::CPB;  // workgroups per instance
  static constexpr int RG = N / R::R;                        // of which run the row passes
  static constexpr int WG = C::WG;
  static constexpr int LDS = RowLds<N>::X + RowLds<N>::TL > ColLds<N>::X ? RowLds<N>::X + RowLds<N>::TL
                                                                         : ColLds<N>::X;
  static_assert(R::WG == C::WG, "row and column passes must share the workgroup size");
  static_assert(RG <= CG, "row groups must fit the instance's workgroups");
};

// Barrier of the workgroups of one instance (see above). Returns
// false when this workgroup gave up or found the launch's failure word set.
__device__ __forceinline__ bool coop_barrier(unsigned* ctr, unsigned target, unsigned* fail, int spin_limit,
                                             int* flag) {
  // the pass's stores stay above the wait (a compiler-level ordering: a
  // wavefront-scope release needs no cache maintenance), and the wait
  // returns once this lane's stores have completed
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    int polls = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++polls > spin_limit || __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_fetch_or(fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

template <int N, int KIND>
__global__ void __launch_bounds__(CoopCfg<N>::WG) coop2d_kernel(Coop2DArgs ca) {
  using CC = CoopCfg<N>;
  __shared__ cf lds[CC::LDS];
  __shared__ int flag;
  const int inst = blockIdx.x / CC::CG, g = blockIdx.x % CC::CG;
  if (inst >= ca.ninst) return;  // the launcher sizes the grid exactly; never taken
  unsigned* ctr = ca.sync + (long long)inst * 64;
  unsigned* fail = ca.sync + (long long)ca.ninst * 64;
  unsigned arrivals = 0;
  auto meet = [&]() {
    arrivals += CC::CG;
    return coop_barrier(ctr, arrivals, fail, ca.spin_limit, &flag);
  };
  Step2DArgs a = ca.s;
  a.u_in = ca.u0;
  a.u_in_bstride = ca.u0_bstride;
  a.u_out_bstride = ca.traj_bstride;
  a.write_T = 0;
  cf* rx = lds;
  cf* rt = lds + RowLds<N>::X;
  const bool row = g < CC::RG;
  const int n = ca.nsteps;
  int next_out = 0;  // first state this workgroup's rows have not been written to yet
  // after a barrier gave up: this workgroup's rows of every state not yet
  // written become NaN, so no stale data passes for a result
  auto give_up = [&]() {
    if (!row) return;
    const int x0 = g * RowCfg<N>::R;
    const int t0 = ca.final_only ? 0 : next_out, t1 = ca.final_only ? 1 : n;
    for (int t = t0; t < t1; ++t) {
      float* o = ca.traj + (long long)inst * ca.traj_bstride + (long long)t * ca.traj_tstride + (long long)x0 * N;
      for (int i = threadIdx.x; i < RowCfg<N>::R * N; i += CC::WG) o[i] = __builtin_nanf("");
    }
  };

  if (row) row2d_body<N, ROW_RA, KIND, true>(a, g, inst, rx, rt);
  if (!meet()) return give_up();
  col2d_body<N, COL_CA, KIND, true>(a, g, inst, lds);
  if (!meet()) return give_up();
  for (int t = 0; t < n; ++t) {
    a.Gout = ca.G ? ca.G + (long long)t * ca.G_tstride : nullptr;
    a.gstage = 0;
    if (row) {
      if (t == 0 || ca.final_only) {
        row2d_body<N, ROW_RB, KIND, true>(a, g, inst, rx, rt);
      } else {
        a.u_out = ca.traj + (long long)(t - 1) * ca.traj_tstride;
        row2d_body<N, ROW_RCB, KIND, true>(a, g, inst, rx, rt);
        next_out = t;
      }
    }
    if (!meet()) return give_up();
    col2d_body<N, COL_CB, KIND, true>(a, g, inst, lds);
    if (!meet()) return give_up();
    a.gstage = 1;
    if (row) row2d_body<N, ROW_RB, KIND, true>(a, g, inst, rx, rt);
    if (!meet()) return give_up();
    a.next_fields = t < n - 1 ? 1 : 0;
    col2d_body<N, COL_CCA, KIND, true>(a, g, inst, lds);
    if (!meet()) return give_up();
  }
  if (row) {
    a.u_out = ca.final_only ? ca.traj : ca.traj + (long long)(n - 1) * ca.traj_tstride;
    row2d_body<N, ROW_RC, KIND, true>(a, g, inst, rx, rt);
  }
}

// The transposed steps of one rollout segment (gradient norm, ETDRK2) as
// one cooperative launch: for i = len - 1 .. 0 the VJP of the step at the
// segment's state i applied to the cotangent acc (in place), the launch
// sequence of the adjoint step (adjoint2d.hpp: RA(u) CA RBG CB RG, then
// RA(g) CL RP CD RP CE RC) with barriers between passes; RG shares a pass
// with RA(g), and a step's RC with the next step's RA(u) (each row
// workgroup reads its tile of T before it writes T). The gradient fields G,
// the cotangent acc and the column states (S0, S1, N0, LAM, RHO) stay with
// the workgroup that wrote them; T, F, H and P are accessed coherently.
// With Gseg (the gradient fields of the segment's steps, stored by the
// forward launch that recomputed the segment) the recompute passes are
// skipped: a step is RC + RA(g), CL, RP, CD, RP, CE (six passes).
struct CoopAdj2DArgs {
  Adj2DArgs aa;
  const float* Gseg;       // the segment's gradient fields from the forward launch (step i at Gseg +
  long long G_tstride;     // i G_tstride), or null: recompute them from the segment's states
  const float* seg;        // the segment's states: state i of instance b at seg + b * seg_bstride + i * seg_tstride
  long long seg_bstride, seg_tstride;
  int len;
  float* acc;              // cotangent (ninst, N, N), in place
  long long acc_bstride;
  int ninst;
  unsigned* sync;
  int spin_limit;
};

template <int N>
__global__ void __launch_bounds__(CoopCfg<N>::WG) coop_adj2d_kernel(CoopAdj2DArgs ca) {
  using CC = CoopCfg<N>;
  constexpr int K = NL_GRADNORM;
  __shared__ cf lds[CC::LDS];
  __shared__ int flag;
  const int inst = blockIdx.x / CC::CG, g = blockIdx.x % CC::CG;
  if (inst >= ca.ninst) return;  // the launcher sizes the grid exactly; never taken
  unsigned* ctr = ca.sync + (long long)inst * 64;
  unsigned* fail = ca.sync + (long long)ca.ninst * 64;
  unsigned arrivals = 0;
  auto meet = [&]() {
    arrivals += CC::CG;
    return coop_barrier(ctr, arrivals, fail, ca.spin_limit, &flag);
  };
  Adj2DArgs aa = ca.aa;
  Step2DArgs& a = aa.s;
  a.write_T = 0;
  cf* rx = lds;
  cf* rt = lds + RowLds<N>::X;
  const bool row = g < CC::RG;
  // after a barrier gave up: NaN in this workgroup's rows of the cotangent
  auto give_up = [&]() {
    if (!row) return;
    float* o = ca.acc + (long long)inst * ca.acc_bstride + (long long)g * RowCfg<N>::R * N;
    for (int i = threadIdx.x; i < RowCfg<N>::R * N; i += CC::WG) o[i] = __builtin_nanf("");
  };
  auto emit = [&]() {  // RC: the transposed step's result into acc
    a.u_out = ca.acc;
    a.u_out_bstride = ca.acc_bstride;
    row2d_body<N, ROW_RC, K, true>(a, g, inst, rx, rt);
  };
  for (int i = ca.len - 1; i >= 0; --i) {
    if (ca.Gseg) {
      // the gradient fields of both stages come from the forward launch:
      // only the transposed stages run (RC of the previous step, RA(g))
      aa.G = const_cast<float*>(ca.Gseg) + (long long)i * ca.G_tstride;
      if (row) {
        if (i < ca.len - 1) {
          emit();
          __syncthreads();
        }
        a.u_in = ca.acc;
        a.u_in_bstride = ca.acc_bstride;
        row2d_body<N, ROW_RA, K, true>(a, g, inst, rx, rt);
      }
      if (!meet()) return give_up();
    } else {
      if (row) {
        if (i < ca.len - 1) {
          emit();
          __syncthreads();
        }
        a.u_in = ca.seg + (long long)i * ca.seg_tstride;
        a.u_in_bstride = ca.seg_bstride;
        row2d_body<N, ROW_RA, K, true>(a, g, inst, rx, rt);
      }
      if (!meet()) return give_up();
      col2d_body<N, COL_CA, K, true>(a, g, inst, lds);
      if (!meet()) return give_up();
      if (row) {
        aa.stage = 0;
        row2d_adj_body<N, ROWA_RBG, true>(aa, g, inst, rx, rt);
      }
      if (!meet()) return give_up();
      col2d_body<N, COL_CB, K, true>(a, g, inst, lds);
      if (!meet()) return give_up();
      if (row) {
        aa.stage = 1;
        row2d_adj_body<N, ROWA_RG, true>(aa, g, inst, rx, rt);
        __syncthreads();
        a.u_in = ca.acc;
        a.u_in_bstride = ca.acc_bstride;
        row2d_body<N, ROW_RA, K, true>(a, g, inst, rx, rt);
      }
      if (!meet()) return give_up();
    }
    col2d_adj_body<N, COLA_CL, true>(aa, g, inst, lds);
    if (!meet()) return give_up();
    if (row) {
      aa.stage = 1;
      row2d_adj_body<N, ROWA_RP, true>(aa, g, inst, rx, rt);
    }
    if (!meet()) return give_up();
    col2d_adj_body<N, COLA_CD, true>(aa, g, inst, lds);
    if (!meet()) return give_up();
    if (row) {
      aa.stage = 0;
      row2d_adj_body<N, ROWA_RP, true>(aa, g, inst, rx, rt);
    }
    if (!meet()) return give_up();
    col2d_adj_body<N, COLA_CE, true>(aa, g, inst, lds);
    if (!meet()) return give_up();
  }
  if (row) emit();
}

template <int N>
__global__ void coop2d_zero_kernel(unsigned* sync, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) sync[i] = 0u;
}

// Instances one launch can hold with every workgroup resident on an
// otherwise idle device. (Work running beside it only delays the launch's
// later workgroups: it does not wait on them. Two cooperative launches that
// together exceed the device could wait on each other; their barriers then
// give up as described above.)
template <int N, int KIND>
inline long long coop2d_max_instances() {
  static long long cached = -1;
  if (cached >= 0) return cached;
  int dev = 0, per_cu = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, coop2d_kernel<N, KIND>, CoopCfg<N>::WG, 0) !=
          hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  cached = (long long)per_cu * cus / CoopCfg<N>::CG;
  return cached;
}

template <int N>
inline long long coop_adj2d_max_instances() {
  static long long cached = -1;
  if (cached >= 0) return cached;
  int dev = 0, per_cu = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, coop_adj2d_kernel<N>, CoopCfg<N>::WG, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  cached = (long long)per_cu * cus / CoopCfg<N>::CG;
  return cached;
}

template <int N>
inline hipError_t launch_coop_adj2d(const CoopAdj2DArgs& a, hipStream_t st) {
  if (a.ninst < 1 || a.len < 1) return hipErrorInvalidValue;
  if (a.ninst > coop_adj2d_max_instances<N>()) return hipErrorCooperativeLaunchTooLarge;
  const int nsync = a.ninst * 64 + 64;
  hipLaunchKernelGGL(coop2d_zero_kernel<N>, dim3((nsync + 255) / 256), dim3(256), 0, st, a.sync, nsync);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((coop_adj2d_kernel<N>), dim3(a.ninst * CoopCfg<N>::CG), dim3(CoopCfg<N>::WG), 0, st, a);
  return hipGetLastError();
}

template <int N, int KIND>
inline hipError_t launch_coop2d(const Coop2DArgs& a, hipStream_t st) {
  if (a.ninst < 1 || a.nsteps < 1) return hipErrorInvalidValue;
  if (a.ninst > coop2d_max_instances<N, KIND>()) return hipErrorCooperativeLaunchTooLarge;
  const int nsync = a.ninst * 64 + 64;
  hipLaunchKernelGGL(coop2d_zero_kernel<N>, dim3((nsync + 255) / 256), dim3(256), 0, st, a.sync, nsync);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((coop2d_kernel<N, KIND>), dim3(a.ninst * CoopCfg<N>::CG), dim3(CoopCfg<N>::WG), 0, st, a);
  return hipGetLastError();
}

}  // namespace exmi
