// FP64 dense hop loop: the reference's arithmetic on the GPU.
//
// Restates _simulate_discrete_record (kmc_dopant_networks.py:33-135) with the
// same FP64 operations in the same order, so that with a replayed uniform
// stream the event sequence equals the CPU restatement's bit for bit:
//   * site energies: se[i] = E[i] + ((-I_0)*R) * sum_{j != i, j empty} 1/d[i][j],
//     summed over ascending j (:59-66); 1/d is precomputed (same IEEE quotient)
//     and an occupied j adds an exact +0.0, so the partial sums are unchanged;
//   * rates: tc[i][j] * (dE > 0 ? nu*exp(-dE/kT) : nu) for possible pairs (:69-87);
//   * prefix sum in row-major order (:91-97).  Impossible pairs contribute an
//     exact 0.0, so the sequential chain runs over the possible pairs only, in
//     row-major order, which leaves every partial sum unchanged;
//   * hop time -log(1-u)*(1/total) and event = first k with problist[k]/total >= u
//     (:100-110); u == 0 selects k = 0 exactly as the reference's scan does.
// This file is compiled with -ffp-contract=off: no a*b+c is fused.
//
// One wave (64 lanes) per trajectory.  The pair loop, rate evaluation and the
// final search are lane-parallel; only the prefix-sum chain is serial, run by
// one lane eight values at a time from registers.  With STAGE the layout's
// tables (1/d, tc, (I_0*R)/d) are copied to LDS once per trajectory.
#include "kmc_device.h"

namespace kmc {

// _transition_possible, kmc_dopant_networks.py:137-163
__device__ __forceinline__ bool possible(int i, int j, int N, const Occ4& o) {
  if (i >= N && j >= N) return false;
  if (i >= N) return !o.test(j);
  if (j >= N) return o.test(i);
  if (i == j) return false;
  return o.test(i) && !o.test(j);
}

// LDS carve-up for S sites, N acceptors (all offsets in doubles, 16-byte aligned).
struct PlainLds {
  size_t se, buf, list, invd, tc, ird, total;
  __host__ __device__ PlainLds(int S, int N, bool stage) {
    auto up2 = [](size_t x) { return (x + 1) & ~(size_t)1; };
    const size_t SS = (size_t)S * S;
    se = 0;
    buf = up2(S);
    list = buf + up2(SS);                       // int16 entries, SS/4 doubles
    invd = list + up2((SS * sizeof(int16_t) + 7) / 8);
    tc = invd + (stage ? up2((size_t)N * N) : 0);
    ird = tc + (stage ? up2(SS) : 0);
    total = ird + (stage ? up2(SS) : 0);
  }
};

template <bool REC, bool STAGE>
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

  double* dl = reinterpret_cast<double*>(smem);
  const PlainLds off(S, N, STAGE);
  double* se = dl + off.se;
  double* buf = dl + off.buf;     // rates of the possible pairs, then their prefix sums
  int16_t* list = reinterpret_cast<int16_t*>(dl + off.list);
  const double* invd = L.invd64;
  const double* tcp = L.tc64;
  const double* irdp = L.ird64;
  if (STAGE) {
    double* a = dl + off.invd;
    double* b = dl + off.tc;
    double* c = dl + off.ird;
    for (int k = lane; k < N * N; k += 64) a[k] = L.invd64[k];
    for (int k = lane; k < SS; k += 64) {
      b[k] = L.tc64[k];
      c[k] = L.ird64[k];
    }
    invd = a;
    tcp = b;
    irdp = c;
  }

  Occ4 occ{td.occ[0], td.occ[1], td.occ[2], td.occ[3]};

  for (int e = lane; e < L.P; e += 64) se[N + e] = es.V[e];
  double E0 = 0.0, E1 = 0.0;
  if (lane < N) E0 = es.E[lane];
  if (lane + 64 < N) E1 = es.E[lane + 64];
  __syncthreads();
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
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const int i = lane + 64 * part;
      if (i < N) {
        double acc = 0.0;
        for (int j = 0; j < N; ++j) {
          const bool add = !occ.test(j) && j != i;
          const double x = invd[j * N + i];
          acc += add ? x : 0.0;
        }
        se[i] = (part ? E1 : E0) + nIR * acc;
      }
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
      if (i < N && j < N) dE = se[j] - se[i] - irdp[k];
      else dE = se[j] - se[i];
      double m;
      if (dE > 0) m = nu * exp(-dE / kT);
      else m = nu;
      buf[idx] = tcp[k] * m;
    }
    __syncthreads();
    // cumulative sum, :91-97: one lane, eight values per LDS round trip
    double acc = 0.0;
    if (lane == 0) {
      int idx = 0;
      for (; idx + 8 <= n; idx += 8) {
        double r[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] = buf[idx + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          acc = r[q] + acc;
          r[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) buf[idx + q] = r[q];
      }
      for (; idx < n; ++idx) {
        acc = buf[idx] + acc;
        buf[idx] = acc;
      }
    }
    const double total = readlane_f64(acc, 0);
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
        const bool hit = idx < n && buf[idx] / total >= ue;
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
    if (from < N) occ.assign(from, false);
    else if (lane == from - N) eo -= 1;
    if (to < N) occ.assign(to, true);
    else if (lane == to - N) eo += 1;
    // record, :126-130
    if (REC && tally) {
      if (p.traffic && lane == 0) p.traffic[(size_t)t * SS + k_ev] += 1.0;
      if (lane < N && occ.test(lane)) ot0 += hop_time;
      if (lane + 64 < N && occ.test(lane + 64)) ot1 += hop_time;
    }
    time += hop_time;
    __syncthreads();
  }
  TrajOut* o = p.out + t;
  if (lane == 0) {
    o->time = time;
    o->hops_done = h > td.prehops ? h - td.prehops : 0;
    o->status = status;
    o->occ[0] = occ.a;
    o->occ[1] = occ.b;
    o->occ[2] = occ.c;
    o->occ[3] = occ.d;
  }
  if (lane < L.P) o->eo[lane] = eo;
  if (REC && p.occ_time) {
    if (lane < N) p.occ_time[(size_t)t * NMAX + lane] = ot0;
    if (lane + 64 < N) p.occ_time[(size_t)t * NMAX + lane + 64] = ot1;
  }
  if (p.se_out)
    for (int i = lane; i < N; i += 64) p.se_out[(size_t)t * NMAX + i] = se[i];
}

template __global__ void kmc_plain_fp64_kernel<false, false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<true, false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<false, true>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<true, true>(PlainParams);

// LDS bytes the kernel needs for S sites, N acceptors.
size_t plain_lds_bytes(int S, int N, bool stage) {
  return PlainLds(S, N, stage).total * sizeof(double);
}

}  // namespace kmc
