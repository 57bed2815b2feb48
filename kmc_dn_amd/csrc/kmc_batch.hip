// Solo kernels of the FP32 batched path: one wave per trajectory.
//
// Workgroups (table policies) or waves (positions) pull work from a global
// atomic queue until it is drained, so one launch serves uneven batches.  The
// per-trajectory hop loop is run_trajectory (kmc_traj.h).
#include "kmc_traj.h"

namespace kmc {

// Table policies: the workgroup's W waves run W trajectories of one layout and
// share its tables in LDS.  Layout: the waves' state stores (W *
// store_wave_bytes(p.store_slots)), the waves' electrode factors (W*2*PMAX
// floats), the electrode weights (elw_lds_floats), the pair table
// (S*S + 64C float2) and, for BT, the c table (N*S + 64C floats).  PH: the
// phase-timing instantiation (run_trajectory).
template <int C, int W, bool REC, bool BT, bool PRIO, bool PH>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(C == 1 ? 4 : 1)))
kmc_batch_fp32_kernel(BatchParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  // wave index and everything read from the shared group slot are wave-uniform:
  // readfirstlane lets the compiler keep them (and the trajectory state derived
  // from them) in scalar registers with scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int sb = store_wave_bytes(p.store_slots);
  unsigned char* store = sb ? smem + wave * sb : nullptr;
  float* evc = reinterpret_cast<float*>(smem + W * sb) + wave * 2 * PMAX;
  float* elw = reinterpret_cast<float*>(smem + W * sb) + W * 2 * PMAX;
  __shared__ int s_group;
  __shared__ int s_layout;
  if (threadIdx.x == 0) s_layout = -1;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_group = atomicAdd(p.queue, 1);
    __syncthreads();
    const int g = __builtin_amdgcn_readfirstlane(s_group);
    if (g >= p.ngroups) break;
    const GroupDev gd = p.groups[g];
    const LayoutDev L = p.layouts[gd.layout];
    // a network of S <= 64 sites runs in one-chunk code (its tables laid out
    // for one chunk): the bits of any chunk count, with the state store
    const bool one = C > 1 && L.S <= 64;
    const int Ce = one ? 1 : C;
    float2* tab = reinterpret_cast<float2*>(elw + elw_lds_floats(L.P, Ce));
    float* ctab = reinterpret_cast<float*>(tab + L.S * L.S + 64 * Ce);
    if (gd.layout != __builtin_amdgcn_readfirstlane(s_layout)) {
      if (one) load_elw<1>(elw, L, threadIdx.x, 64 * W);
      else load_elw<C>(elw, L, threadIdx.x, 64 * W);
      const int n = L.S * L.S + 64 * Ce;
      const float2* src = BT ? L.tabB : L.tab32;
      for (int k = threadIdx.x; k < n; k += 64 * W) tab[k] = src[k];
      if (BT) {
        const int nc = L.N * L.S + 64 * Ce;
        for (int k = threadIdx.x; k < nc; k += 64 * W) ctab[k] = L.ccol[k];
      }
      __syncthreads();
      if (threadIdx.x == 0) s_layout = gd.layout;
    }
    __syncthreads();
    if (wave < gd.count) {
      if (one) {
        const TabPolicy<1, BT> pol = TabPolicy<1, BT>::make(tab, ctab, elw, evc, L.S, lane);
        run_trajectory<1, REC, TabPolicy<1, BT>, 1, PRIO, PH, false>(p, gd.first + wave, L, pol,
                                                                     lane, Team<1, 1>{}, store);
      } else {
        const TabPolicy<C, BT> pol = TabPolicy<C, BT>::make(tab, ctab, elw, evc, L.S, lane);
        run_trajectory<C, REC, TabPolicy<C, BT>, 1, PRIO, PH>(p, gd.first + wave, L, pol, lane,
                                                              Team<1, C>{}, store);
      }
    }
  }
}

// Positions policy: every wave pulls its own groups and keeps only its layout's
// site coordinates in LDS (at most 64*C float4), so a workgroup mixes layouts.
// The waves' state stores follow the coordinates (W * store_wave_bytes).
template <int C, int W, bool REC>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(C == 1 ? 4 : 1)))
kmc_batch_pos_kernel(BatchParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  float4* pos = reinterpret_cast<float4*>(smem) + wave * (64 * C);
  const int sb = store_wave_bytes(p.store_slots);
  unsigned char* store = sb ? smem + W * 64 * C * sizeof(float4) + wave * sb : nullptr;
  for (;;) {
    int g = 0;
    if (lane == 0) g = atomicAdd(p.queue, 1);
    g = __builtin_amdgcn_readfirstlane(readlane_i32(g, 0));
    if (g >= p.ngroups) break;
    const GroupDev gd = p.groups[g];
    const LayoutDev L = p.layouts[gd.layout];
    for (int k = lane; k < L.S; k += 64) pos[k] = L.pos4[k];
    wave_lds_sync();
    if (C > 1 && L.S <= 64) {
      // a one-chunk network: the one-chunk code (fast hops and leaps, fewer
      // instructions per hop); its trajectories' bits are those of any chunk count
      PosPolicy<1> pol;
      pol.init(pos, L, lane);
      for (int i = 0; i < gd.count; ++i)
        run_trajectory<1, REC, PosPolicy<1>, 1, true, false, false>(p, gd.first + i, L, pol, lane,
                                                                    Team<1, 1>{}, store);
    } else {
      PosPolicy<C> pol;
      pol.init(pos, L, lane);
      for (int i = 0; i < gd.count; ++i)
        run_trajectory<C, REC, PosPolicy<C>, 1, true>(p, gd.first + i, L, pol, lane,
                                                      Team<1, C>{}, store);
    }
    wave_lds_sync();
  }
}

#define KMC_INST(C, W)                                                                           \
  template __global__ void kmc_batch_fp32_kernel<C, W, false, false, true, false>(BatchParams);  \
  template __global__ void kmc_batch_fp32_kernel<C, W, true, false, true, false>(BatchParams);   \
  template __global__ void kmc_batch_fp32_kernel<C, W, false, true, true, false>(BatchParams);   \
  template __global__ void kmc_batch_fp32_kernel<C, W, true, true, true, false>(BatchParams);    \
  template __global__ void kmc_batch_fp32_kernel<C, W, false, false, false, false>(BatchParams); \
  template __global__ void kmc_batch_fp32_kernel<C, W, false, true, false, false>(BatchParams);  \
  template __global__ void kmc_batch_pos_kernel<C, W, false>(BatchParams);                       \
  template __global__ void kmc_batch_pos_kernel<C, W, true>(BatchParams);
KMC_INST(1, 1) KMC_INST(1, 2) KMC_INST(1, 4) KMC_INST(1, 8)
KMC_INST(2, 1) KMC_INST(2, 2) KMC_INST(2, 4) KMC_INST(2, 8)
KMC_INST(3, 1) KMC_INST(3, 2) KMC_INST(3, 4) KMC_INST(3, 8)
#undef KMC_INST
// phase timing: 4 waves per workgroup, both table forms, graded issue priority
#define KMC_INST(C)                                                                           \
  template __global__ void kmc_batch_fp32_kernel<C, 4, false, false, true, true>(BatchParams); \
  template __global__ void kmc_batch_fp32_kernel<C, 4, false, true, true, true>(BatchParams);
KMC_INST(1) KMC_INST(2) KMC_INST(3)
#undef KMC_INST

// Zeroes the work-queue head (a kernel node, so a captured graph can replay it).
__global__ void kmc_queue_reset_kernel(int32_t* q) { *q = 0; }

// Zeroes n doubles (grid-stride): clears record buffers on the stream, also
// inside a captured graph, where a memset node is not used.
__global__ void kmc_fill_zero_kernel(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Scatters n staged energy sets into their device slots, one workgroup per set:
// set i (id ids[i], nps[2i] acceptors, nps[2i+1] electrodes) takes E[i*NMAX ...]
// and V[i*PMAX ...].
__global__ void kmc_eset_scatter_kernel(const EsetDev* esets, const int32_t* ids, const int32_t* nps,
                                        const double* E, const double* V, int64_t n) {
  const int64_t i = blockIdx.x;
  if (i >= n) return;
  const EsetDev e = esets[ids[i]];
  const int N = nps[2 * i], P = nps[2 * i + 1];
  double* Ed = const_cast<double*>(e.E);
  double* Vd = const_cast<double*>(e.V);
  for (int k = threadIdx.x; k < N; k += blockDim.x) Ed[k] = E[i * NMAX + k];
  for (int k = threadIdx.x; k < P; k += blockDim.x) Vd[k] = V[i * PMAX + k];
}

// Upper bound of the LDS bytes of table-form layouts with at most S sites, N
// acceptors and P electrodes, for W waves per workgroup.
size_t batch_lds_bytes(bool bt, int C, int W, int S, int N, int P) {
  size_t b = (size_t)W * 2 * PMAX * sizeof(float);
  b += (size_t)elw_lds_floats(P, C) * sizeof(float);
  b += ((size_t)S * S + 64 * C) * sizeof(float2);
  if (bt) b += ((size_t)N * S + 64 * C) * sizeof(float);
  return b;
}

size_t pos_lds_bytes(int C, int W) { return (size_t)W * 64 * C * sizeof(float4); }

// LDS bytes of the W waves' state stores of `slots` slots each (they come before
// the table kernels' other LDS and after the positions kernels' coordinates).
size_t store_lds_bytes(int W, int slots) { return (size_t)W * store_wave_bytes(slots); }

}  // namespace kmc
