// Team kernels of the FP32 batched path: T waves (one workgroup) per trajectory.
//
// For long trajectories on a GPU that the batch does not fill, the hop latency
// of one trajectory bounds the launch.  A team splits the walked set of every
// hop over T waves on different SIMDs (Team in kmc_traj.h); the remaining
// per-hop work is replicated.  Workgroups pull one trajectory at a time from
// the atomic queue.  FORM: 1 {lt, lt+c} + c table, 2 {lt, c} table, 3 positions;
// table forms keep the layout's table in LDS while consecutive trajectories
// share it.
#include "kmc_traj.h"

namespace kmc {

// PH: the phase-timing instantiation (run_trajectory).
template <int C, int T, bool REC, int FORM, bool PH>
__global__ void __launch_bounds__(64 * T) kmc_team_kernel(BatchParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  Team<T, C> team;
  team.wave = wave;
  team.red = reinterpret_cast<float*>(smem);
  unsigned char* rest = smem + (size_t)2 * T * C * 64 * sizeof(float);
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
    if constexpr (FORM == 3) {
      float4* pos = reinterpret_cast<float4*>(rest);
      for (int k = threadIdx.x; k < L.S; k += 64 * T) pos[k] = L.pos4[k];
      __syncthreads();
      PosPolicy<C> pol;
      pol.init(pos, L, lane);
      for (int i = 0; i < gd.count; ++i)
        run_trajectory<C, REC, PosPolicy<C>, T, true, PH>(p, gd.first + i, L, pol,
                                                                       lane, team);
    } else {
      // the waves' electrode factors (T*2*PMAX floats), the electrode weights,
      // the pair table and, for FORM 1, the c table
      constexpr bool BT = FORM == 1;
      float* evc = reinterpret_cast<float*>(rest) + wave * 2 * PMAX;
      float* elw = reinterpret_cast<float*>(rest) + T * 2 * PMAX;
      float2* tab = reinterpret_cast<float2*>(elw + elw_lds_floats(L.P, C));
      float* ctab = reinterpret_cast<float*>(tab + L.S * L.S + 64 * C);
      if (gd.layout != __builtin_amdgcn_readfirstlane(s_layout)) {
        load_elw<C>(elw, L, threadIdx.x, 64 * T);
        const int n = L.S * L.S + 64 * C;
        const float2* src = BT ? L.tabB : L.tab32;
        for (int k = threadIdx.x; k < n; k += 64 * T) tab[k] = src[k];
        if (BT) {
          const int nc = L.N * L.S + 64 * C;
          for (int k = threadIdx.x; k < nc; k += 64 * T) ctab[k] = L.ccol[k];
        }
        __syncthreads();
        if (threadIdx.x == 0) s_layout = gd.layout;
      }
      __syncthreads();
      const TabPolicy<C, BT> pol = TabPolicy<C, BT>::make(tab, ctab, elw, evc, L.S, lane);
      for (int i = 0; i < gd.count; ++i)
        run_trajectory<C, REC, TabPolicy<C, BT>, T, true, PH>(p, gd.first + i, L,
                                                                           pol, lane, team);
    }
  }
}

#define KMC_TEAM(C, T)                                                           \
  template __global__ void kmc_team_kernel<C, T, false, 1, false>(BatchParams); \
  template __global__ void kmc_team_kernel<C, T, true, 1, false>(BatchParams);  \
  template __global__ void kmc_team_kernel<C, T, false, 2, false>(BatchParams); \
  template __global__ void kmc_team_kernel<C, T, true, 2, false>(BatchParams);  \
  template __global__ void kmc_team_kernel<C, T, false, 3, false>(BatchParams); \
  template __global__ void kmc_team_kernel<C, T, true, 3, false>(BatchParams);
KMC_TEAM(1, 2) KMC_TEAM(1, 4) KMC_TEAM(2, 2) KMC_TEAM(2, 4) KMC_TEAM(3, 2) KMC_TEAM(3, 4)
#undef KMC_TEAM
// phase timing: teams of 4 on two- and three-chunk networks
#define KMC_TEAM(C, F) template __global__ void kmc_team_kernel<C, 4, false, F, true>(BatchParams);
KMC_TEAM(2, 1) KMC_TEAM(2, 2) KMC_TEAM(2, 3) KMC_TEAM(3, 1) KMC_TEAM(3, 2) KMC_TEAM(3, 3)
#undef KMC_TEAM

// LDS bytes of a team workgroup: partial row sums plus the form's tables or
// positions (an upper bound for layouts with at most S, N and P).
size_t team_lds_bytes(int form, int C, int T, int S, int N, int P) {
  size_t b = (size_t)2 * T * C * 64 * sizeof(float);
  if (form == 3) return b + (size_t)64 * C * sizeof(float4);
  b += (size_t)T * 2 * PMAX * sizeof(float);
  b += (size_t)elw_lds_floats(P, C) * sizeof(float);
  b += ((size_t)S * S + 64 * C) * sizeof(float2);
  if (form == 1) b += ((size_t)N * S + 64 * C) * sizeof(float);
  return b;
}

}  // namespace kmc
