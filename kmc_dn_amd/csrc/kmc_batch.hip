// FP32-rate / FP64-accumulate batched hop loop (one wave per trajectory).
//
// Same model as _simulate_discrete_record (kmc_dopant_networks.py:33-135) and
// Go's simulate (goSimulation/simulation.go:194-325), organised for CDNA4:
//
//  * Each lane owns sites s = lane + 64*c (c < C).  Site energies live in
//    registers as e~ = se * log2(e)/kT; electrodes hold V * log2(e)/kT.
//  * The pair table of the group's layout sits in LDS, shared by the W waves of
//    the workgroup, which run W trajectories of that layout.  With
//    lt = log2(nu * tc) and c = scaled Coulomb term I_0*R/d, a rate is
//        2^(lt + min(0, e~_src - e~_dst + c)) = 2^min(lt, b + e~_src - e~_dst),
//    b = lt + c: one v_exp_f32 per pair and no FP32 underflow of the 10^-50
//    prefactors.  Table form BT: {lt, b} pairs plus a c table for the energy
//    update (fewest VALU ops); form !BT: {lt, c} pairs only (smallest LDS,
//    for the largest networks).
//  * Per hop, the smaller of {sources} (occupied acceptors + electrodes) and
//    {targets} (empty acceptors + electrodes) is walked four members at a time
//    with a scalar loop while the lanes cover every site of the other role:
//    row sums R_lane accumulate in registers with no cross-lane traffic.
//  * Event pick: FP64 wave prefix sum of R (DPP), Philox event uniform * total,
//    ballot for the first lane whose prefix exceeds it, then one lane-parallel
//    pass over that lane's pairs (FP32 DPP scan) for the partner site.
//  * Hop: occupation bits flip in SGPRs; energies get the O(N) rank-1 Coulomb
//    update of makeJump (simulation.go:107-130); electrode tallies stay in the
//    electrode lanes' registers; the clock is accumulated in FP64.
//  * Every `recompute_interval` hops the energies are rebuilt in FP64 from
//    E_constant and the I_0*R/d table, cancelling FP32 drift of the updates
//    (the Go reference never re-synchronises).
//  * Uniforms come from one Philox block per lane every 64 hops (lane k holds
//    hop h0+k), so the RNG costs two readlanes per hop.
//
// The workgroup pulls groups (layout, first trajectory, count) from a global
// atomic queue until it is drained, so one launch serves uneven batches.
#include "kmc_device.h"

namespace kmc {

template <int C>
__device__ __forceinline__ void recompute_energies(float (&se)[C], const uint64_t (&occ)[OCCW],
                                                   const LayoutDev& L, const EsetDev& es,
                                                   int lane) {
  const int N = L.N, S = L.S;
  const double sc = (double)L.scale2;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    float v = 0.f;
    if (s < N) {
      double acc = 0.0;
      for (int j = 0; j < N; ++j)
        if (!((occ[j >> 6] >> (j & 63)) & 1ull) && j != s) acc += L.coul64[j * N + s];
      v = (float)((es.E[s] - acc) * sc);
    } else if (s < S) {
      v = (float)(es.V[s - N] * sc);
    }
    se[c] = v;
  }
}

// log2 of a pair rate from its table entry and the source-minus-target energy.
template <bool BT>
__device__ __forceinline__ float log2_rate(float2 tv, float diff) {
  if constexpr (BT) return fminf(tv.x, tv.y + diff);
  else return tv.x + fminf(0.f, diff + tv.y);
}

// Row sums over the uniform set U (bits of umask), lanes covering the other role.
// UDST: U = targets, lanes = sources; otherwise U = sources, lanes = targets.
// Four members of U per step so their LDS reads are in flight together; a
// missing member reads the all -inf column past the table (index S) and adds 0.
template <int C, bool UDST, bool BT>
__device__ __forceinline__ void row_sums(const float2* __restrict__ tab, int S,
                                         const uint64_t (&umask)[C], const float (&v)[C],
                                         const float (&se)[C], int lane, float (&R)[C]) {
  constexpr int U = 4;
#pragma unroll
  for (int cu = 0; cu < C; ++cu) {
    uint64_t m = umask[cu];
    while (m) {
      int b[U], col[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (m) {
          b[k] = __builtin_ctzll(m);
          m &= m - 1;
          col[k] = (64 * cu + b[k]) * S;
        } else {
          b[k] = 0;
          col[k] = S * S;
        }
      }
      float2 tv[U][C];
#pragma unroll
      for (int k = 0; k < U; ++k)
#pragma unroll
        for (int c = 0; c < C; ++c) tv[k][c] = tab[col[k] + lane + 64 * c];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const float su = readlane_f32(se[cu], b[k]);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const float diff = UDST ? (v[c] - su) : (su - v[c]);
          R[c] += fast_exp2(log2_rate<BT>(tv[k][c], diff));
        }
      }
    }
  }
}

template <int C>
__device__ __forceinline__ float pick_lane_value(const float (&x)[C], int site) {
  const int cs = site >> 6, ls = site & 63;
  float r = readlane_f32(x[0], ls);
#pragma unroll
  for (int c = 1; c < C; ++c)
    if (cs == c) r = readlane_f32(x[c], ls);
  return r;
}

template <int C, bool REC, bool BT>
__device__ void run_trajectory(const BatchParams& p, int t, const LayoutDev& L,
                               const float2* __restrict__ tab, const float* __restrict__ ctab,
                               int lane) {
  const TrajDev td = p.trajs[t];
  const EsetDev es = p.esets[td.eset];
  const int N = L.N, S = L.S;
  uint64_t occ[OCCW];
#pragma unroll
  for (int w = 0; w < OCCW; ++w) occ[w] = td.occ[w];

  float se[C];
  int eo[C];
  double ot[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    eo[c] = 0;
    ot[c] = 0.0;
  }
  recompute_energies<C>(se, occ, L, es, lane);

  double time = 0.0;
  int32_t status = ST_OK;
  const int64_t total_hops = td.prehops + td.hops;
  const int K = p.recompute_interval;
  int since = 0;
  float lg_lane = 0.f;    // log2(1 - u_clock) of hop (h & ~63) + lane
  double ue_lane = 0.0;   // event uniform of that hop
  int64_t h = 0;
  for (; h < total_hops; ++h) {
    if ((h & 63) == 0) {
      double uc;
      stream_uniforms(td.seed, td.hop0 + (uint64_t)(h + lane), uc, ue_lane);
      lg_lane = fast_log2((float)(1.0 - uc));
    }
    if (h == td.prehops) {
      time = 0.0;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        eo[c] = 0;
        ot[c] = 0.0;
      }
    }
    if (since == K) {
      recompute_energies<C>(se, occ, L, es, lane);
      since = 0;
    }
    ++since;

    // roles of the lanes' sites
    uint64_t srcm[C], dstm[C];
    bool occd[C];
    int nsrc = 0, ndst = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      const bool isacc = s < N;
      const bool isel = s >= N && s < S;
      occd[c] = isacc && ((occ[s >> 6] >> (s & 63)) & 1ull);
      srcm[c] = ballot(isel || occd[c]);
      dstm[c] = ballot(isel || (isacc && !occd[c]));
      nsrc += __builtin_popcountll(srcm[c]);
      ndst += __builtin_popcountll(dstm[c]);
    }
    const bool udst = ndst <= nsrc;   // walk the smaller set
    float v[C], R[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      R[c] = 0.f;
      const bool lane_ok = udst ? ((srcm[c] >> lane) & 1ull) : ((dstm[c] >> lane) & 1ull);
      v[c] = lane_ok ? se[c] : (udst ? -INFINITY : INFINITY);
    }
    if (udst) row_sums<C, true, BT>(tab, S, dstm, v, se, lane, R);
    else row_sums<C, false, BT>(tab, S, srcm, v, se, lane, R);

    // FP64 prefix over lane rows
    double incl[C];
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      incl[c] = wave_scan_f64((double)R[c]) + total;
      total = readlane_f64(incl[c], 63);
    }
    if (!(total > 0.0) || !(total < INFINITY)) {
      status = ST_ZERO_RATE;
      break;
    }
    const float lg = readlane_f32(lg_lane, (int)(h & 63));
    const double ue = readlane_f64(ue_lane, (int)(h & 63));
    // exponential waiting time -ln(1-u)/total, FP32 quotient into the FP64 clock
    const double dt = (double)(lg * -0.69314718055994531f * __builtin_amdgcn_rcpf((float)total));
    const double target = ue * total;

    int sel = -1;
    double excl = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const uint64_t m = ballot(incl[c] > target);
      if (sel < 0 && m) {
        const int l = __builtin_ctzll(m);
        sel = 64 * c + l;
        excl = readlane_f64(incl[c] - (double)R[c], l);
      }
    }
    if (sel < 0) {   // target rounded onto the total: last lane with a nonzero row
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const uint64_t m = ballot(R[c] > 0.f);
        if (m) {
          const int l = 63 - __builtin_clzll(m);
          sel = 64 * c + l;
          excl = readlane_f64(incl[c] - (double)R[c], l);
        }
      }
    }
    const float rr = (float)(target - excl);

    // partner of `sel`: lane-parallel pass over its row, in site order
    const float s_sel = pick_lane_value<C>(se, sel);
    const float2* row = tab + sel * S + lane;
    float run = 0.f;
    int pick = -1, last_nz = -1;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const bool inU = udst ? ((dstm[c] >> lane) & 1ull) : ((srcm[c] >> lane) & 1ull);
      const float diff = udst ? (s_sel - se[c]) : (se[c] - s_sel);
      float r = fast_exp2(log2_rate<BT>(row[64 * c], diff));
      r = inU ? r : 0.f;
      const float cum = wave_scan_f32(r) + run;
      const uint64_t nz = ballot(r > 0.f);
      const uint64_t hit = ballot(r > 0.f && cum > rr);
      if (pick < 0) {
        if (nz) last_nz = 64 * c + 63 - __builtin_clzll(nz);
        if (hit) pick = 64 * c + __builtin_ctzll(hit);
      }
      run = readlane_f32(cum, 63);
    }
    if (pick < 0) pick = last_nz;
    if (pick < 0) {
      status = ST_ZERO_RATE;
      break;
    }
    const int src = udst ? sel : pick;
    const int dst = udst ? pick : sel;
    const bool tally = h >= td.prehops;

    if (REC && tally && !p.post_hop_record) {
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (occd[c]) ot[c] += dt;
    }
    if (p.traffic && tally && lane == 0) {   // net traffic, Go convention (simulation.go:310-311)
      double* tr = p.traffic + (size_t)t * p.traffic_stride;
      tr[src * S + dst] += 1.0;
      tr[dst * S + src] -= 1.0;
    }
    // perform hop: occupation, rank-1 energy update, tallies
    if (src < N) {
      occ[src >> 6] &= ~(1ull << (src & 63));
#pragma unroll
      for (int c = 0; c < C; ++c)
        se[c] -= BT ? ctab[src * S + lane + 64 * c] : tab[src * S + lane + 64 * c].y;
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (lane + 64 * c == src) eo[c] -= 1;
    }
    if (dst < N) {
      occ[dst >> 6] |= 1ull << (dst & 63);
#pragma unroll
      for (int c = 0; c < C; ++c)
        se[c] += BT ? ctab[dst * S + lane + 64 * c] : tab[dst * S + lane + 64 * c].y;
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (lane + 64 * c == dst) eo[c] += 1;
    }
    if (REC && tally && p.post_hop_record) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int s = lane + 64 * c;
        if (s < N && ((occ[s >> 6] >> (s & 63)) & 1ull)) ot[c] += dt;
      }
    }
    time += dt;
  }

  TrajOut* o = p.out + t;
  if (lane == 0) {
    o->time = time;
    o->hops_done = h > td.prehops ? h - td.prehops : 0;
    o->status = status;
#pragma unroll
    for (int w = 0; w < OCCW; ++w) o->occ[w] = occ[w];
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    if (s >= N && s < S) o->eo[s - N] = eo[c];
    if (REC && p.occ_time && s < N) p.occ_time[(size_t)t * NMAX + s] = ot[c];
  }
}

// LDS: the group layout's pair table (S*S + 192 float2) and, for BT, its c
// table (N*S + 192 floats).
template <int C, int W, bool REC, bool BT>
__global__ void __launch_bounds__(64 * W) kmc_batch_fp32_kernel(BatchParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float2* tab = reinterpret_cast<float2*>(smem);
  __shared__ int s_group;
  __shared__ int s_layout;
  if (threadIdx.x == 0) s_layout = -1;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_group = atomicAdd(p.queue, 1);
    __syncthreads();
    const int g = s_group;
    if (g >= p.ngroups) break;
    const GroupDev gd = p.groups[g];
    const LayoutDev L = p.layouts[gd.layout];
    float* ctab = reinterpret_cast<float*>(tab + L.S * L.S + 64 * 3);
    if (gd.layout != s_layout) {
      const int n = L.S * L.S + 64 * 3;
      const float2* src = BT ? L.tabB : L.tab32;
      for (int k = threadIdx.x; k < n; k += 64 * W) tab[k] = src[k];
      if (BT) {
        const int nc = L.N * L.S + 64 * 3;
        for (int k = threadIdx.x; k < nc; k += 64 * W) ctab[k] = L.ccol[k];
      }
      __syncthreads();
      if (threadIdx.x == 0) s_layout = gd.layout;
    }
    __syncthreads();
    if (wave < gd.count) run_trajectory<C, REC, BT>(p, gd.first + wave, L, tab, ctab, lane);
  }
}

#define KMC_INST(C, W)                                                      \
  template __global__ void kmc_batch_fp32_kernel<C, W, false, false>(BatchParams); \
  template __global__ void kmc_batch_fp32_kernel<C, W, true, false>(BatchParams);  \
  template __global__ void kmc_batch_fp32_kernel<C, W, false, true>(BatchParams);  \
  template __global__ void kmc_batch_fp32_kernel<C, W, true, true>(BatchParams);
KMC_INST(1, 1) KMC_INST(1, 2) KMC_INST(1, 4) KMC_INST(1, 8)
KMC_INST(2, 1) KMC_INST(2, 2) KMC_INST(2, 4) KMC_INST(2, 8)
KMC_INST(3, 1) KMC_INST(3, 2) KMC_INST(3, 4) KMC_INST(3, 8)
#undef KMC_INST

// Zeroes the work-queue head (a kernel node, so a captured graph can replay it).
__global__ void kmc_queue_reset_kernel(int32_t* q) { *q = 0; }

size_t batch_lds_bytes(bool bt, int S, int N) {
  size_t b = ((size_t)S * S + 64 * 3) * sizeof(float2);
  if (bt) b += ((size_t)N * S + 64 * 3) * sizeof(float);
  return b;
}

}  // namespace kmc
