// FP64 dense hop loop: the reference's arithmetic on the GPU.
//
// Restates _simulate_discrete_record (kmc_dopant_networks.py:33-135) with the
// same FP64 operations in the same order, so that with a replayed uniform
// stream the event sequence equals the CPU restatement's bit for bit:
//   * site energies: se[i] = E[i] + ((-I_0)*R) * sum_{j != i, j empty} 1/d[i][j],
//     summed over ascending j (:59-66); 1/d is precomputed (same IEEE quotient);
//   * rates: tc[i][j] * (dE > 0 ? nu*exp(-dE/kT) : nu) for possible pairs (:69-87);
//   * prefix sum in row-major order (:91-97).  Impossible pairs contribute an
//     exact 0.0, so the sequential chain runs over the possible pairs only, in
//     row-major order, which leaves every partial sum unchanged;
//   * hop time -log(1-u)*(1/total) and event = first k with problist[k]/total >= u
//     (:100-110); u == 0 selects k = 0 exactly as the reference's scan does.
// This file is compiled with -ffp-contract=off: no a*b+c is fused.
//
// One wave (64 lanes) per trajectory.  The pair loop, rate evaluation and the
// final search are lane-parallel; only the prefix-sum chain is serial.
#include "kmc_device.h"

namespace kmc {

// Occupation bit of a lane-varying site index (selects among the uniform words).
__device__ __forceinline__ bool occv(const uint64_t (&o)[OCCW], int i) {
  const int w = i >> 6;
  uint64_t x = o[0];
  x = (w == 1) ? o[1] : x;
  x = (w == 2) ? o[2] : x;
  x = (w == 3) ? o[3] : x;
  return (x >> (i & 63)) & 1ull;
}

// _transition_possible, kmc_dopant_networks.py:137-163
__device__ __forceinline__ bool possible(int i, int j, int N, const uint64_t (&o)[OCCW]) {
  if (i >= N && j >= N) return false;
  if (i >= N) return !occv(o, j);
  if (j >= N) return occv(o, i);
  if (i == j) return false;
  return occv(o, i) && !occv(o, j);
}

template <bool REC>
__global__ void __launch_bounds__(64) kmc_plain_fp64_kernel(PlainParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.x;
  const int lane = threadIdx.x;
  const TrajDev td = p.trajs[t];
  const LayoutDev L = p.layouts[td.layout];
  const EsetDev es = p.esets[td.eset];
  const int N = L.N, S = L.S, SS = S * S;
  const double nu = L.nu, kT = L.kT;
  const double nIR = (-L.I0) * L.R;

  double* se = reinterpret_cast<double*>(smem);           // [S]
  double* rates = se + ((S + 1) & ~1);                    // [SS]
  double* pref = rates + SS;                              // [SS]
  int16_t* list = reinterpret_cast<int16_t*>(pref + SS);  // [SS]

  uint64_t occ[OCCW];
#pragma unroll
  for (int w = 0; w < OCCW; ++w) occ[w] = td.occ[w];

  for (int e = lane; e < L.P; e += 64) se[N + e] = es.V[e];
  int64_t eo = 0;                 // lane e < P holds electrode e's tally
  double ot0 = 0.0, ot1 = 0.0;    // post-hop occupation time, sites lane and lane+64
  double time = 0.0;
  int32_t status = ST_OK;
  const int64_t total_hops = td.prehops + td.hops;
  int64_t h = 0;
  for (; h < total_hops; ++h) {
    const bool tally = h >= td.prehops;
    if (h == td.prehops) {   // reset() between prehops and hops (:562-585)
      time = 0.0;
      eo = 0;
      ot0 = ot1 = 0.0;
    }
    // site energies, :59-66
    for (int i = lane; i < N; i += 64) {
      double acc = 0.0;
      for (int j = 0; j < N; ++j)
        if (!((occ[j >> 6] >> (j & 63)) & 1ull) && j != i) acc += L.invd64[j * N + i];
      se[i] = es.E[i] + nIR * acc;
    }
    __syncthreads();
    // possible pairs, compacted in row-major order
    int n = 0;
    for (int k0 = 0; k0 < SS; k0 += 64) {
      const int k = k0 + lane;
      const int i = k / S, j = k - (k / S) * S;
      const bool ok = k < SS && possible(i, j, N, occ);
      const uint64_t m = ballot(ok);
      if (ok) list[n + mbcnt(m)] = (int16_t)k;
      n += __builtin_popcountll(m);
    }
    __syncthreads();
    // rates of the possible pairs, :69-87
    for (int idx = lane; idx < n; idx += 64) {
      const int k = list[idx];
      const int i = k / S, j = k - (k / S) * S;
      double dE;
      if (i < N && j < N) dE = se[j] - se[i] - L.ird64[k];
      else dE = se[j] - se[i];
      double m;
      if (dE > 0) m = nu * exp(-dE / kT);
      else m = nu;
      rates[idx] = L.tc64[k] * m;
    }
    __syncthreads();
    // cumulative sum, :91-97 (serial; every lane runs the same chain)
    double acc = 0.0;
    for (int idx = 0; idx < n; ++idx) {
      acc = rates[idx] + acc;
      if (lane == 0) pref[idx] = acc;
    }
    const double total = acc;
    if (p.totals && t == 0 && lane == 0) p.totals[h] = total;
    if (total == 0.0) {   // reference: ZeroDivisionError at 1/problist[-1]
      status = ST_ZERO_RATE;
      break;
    }
    double uc, ue;
    if (p.uniforms) {
      uc = p.uniforms[2 * h];
      ue = p.uniforms[2 * h + 1];
    } else {
      stream_uniforms(td.seed, td.hop0 + (uint64_t)h, uc, ue);
    }
    const double scale = 1.0 / total;
    const double hop_time = -log(1.0 - uc) * scale;
    __syncthreads();
    // first k with problist[k]/total >= u, :103-110
    int k_ev = 0;
    if (ue != 0.0) {
      int found = -1;
      for (int b = 0; b < n; b += 64) {
        const int idx = b + lane;
        const bool hit = idx < n && pref[idx] / total >= ue;
        const uint64_t m = ballot(hit);
        if (m) {
          found = b + __builtin_ctzll(m);
          break;
        }
      }
      k_ev = found >= 0 ? (int)list[found] : 0;
    }
    const int from = k_ev / S, to = k_ev - (k_ev / S) * S;
    if (p.events && t == 0 && lane == 0) p.events[h] = k_ev;
    // perform hop, :116-123
    if (from < N) occ[from >> 6] &= ~(1ull << (from & 63));
    else if (lane == from - N) eo -= 1;
    if (to < N) occ[to >> 6] |= 1ull << (to & 63);
    else if (lane == to - N) eo += 1;
    // record, :126-130
    if (REC && tally) {
      if (p.traffic && lane == 0) p.traffic[(size_t)t * SS + k_ev] += 1.0;
      if (lane < N && ((occ[lane >> 6] >> (lane & 63)) & 1ull)) ot0 += hop_time;
      if (lane + 64 < N && ((occ[(lane + 64) >> 6] >> ((lane + 64) & 63)) & 1ull)) ot1 += hop_time;
    }
    time += hop_time;
    __syncthreads();
  }
  TrajOut* o = p.out + t;
  if (lane == 0) {
    o->time = time;
    o->hops_done = h > td.prehops ? h - td.prehops : 0;
    o->status = status;
#pragma unroll
    for (int w = 0; w < OCCW; ++w) o->occ[w] = occ[w];
  }
  if (lane < L.P) o->eo[lane] = eo;
  if (REC && p.occ_time) {
    if (lane < N) p.occ_time[(size_t)t * NMAX + lane] = ot0;
    if (lane + 64 < N) p.occ_time[(size_t)t * NMAX + lane + 64] = ot1;
  }
  if (p.se_out)
    for (int i = lane; i < N; i += 64) p.se_out[(size_t)t * NMAX + i] = se[i];
}

template __global__ void kmc_plain_fp64_kernel<false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<true>(PlainParams);

// LDS bytes the kernel needs for S sites.
size_t plain_lds_bytes(int S) {
  const size_t SS = (size_t)S * S;
  return sizeof(double) * (((S + 1) & ~1) + 2 * SS) + sizeof(int16_t) * SS;
}

}  // namespace kmc
