// FP64 dense hop loop: the reference's arithmetic on the GPU.
//
// Restates _simulate_discrete_record (kmc_dopant_networks.py:33-135) with the
// same FP64 operations in the same order, so that with a replayed uniform
// stream the event sequence equals the CPU restatement's bit for bit:
//   * site energies: se[i] = E[i] + ((-I_0)*R) * sum_{j != i} (1 - n_j)/d[i][j],
//     summed over ascending j (:59-66).  An occupied j adds an exact +0.0 to a
//     sum that is +0.0 or positive, which leaves it unchanged, so only the
//     empty acceptors are visited (ascending); 1/d is precomputed (the same
//     IEEE quotient) with a zero diagonal;
//   * rates: tc[i][j] * (dE > 0 ? nu*exp(-dE/kT) : nu) for possible pairs (:69-87);
//   * prefix sum in row-major order (:91-97).  Impossible pairs contribute an
//     exact 0.0, so the sequential chain runs over the possible pairs only, in
//     row-major order, which leaves every partial sum unchanged;
//   * hop time -log(1-u)*(1/total) and event = first k with problist[k]/total >= u
//     (:100-110); u == 0 selects k = 0 exactly as the reference's scan does.
// This file is compiled with -ffp-contract=off: no a*b+c is fused.
//
// One wave (64 lanes) per trajectory.  The possible pairs of a hop have a
// closed structure (_transition_possible, :137-163): rows of occupied acceptors
// (ascending) hop into [empty acceptors ascending, then the electrodes], and
// then each electrode row hops into the empty acceptors.  Two short LDS lists
// (occupied acceptors; empty acceptors followed by electrodes) built with
// ballots give every list position its (row, column) by one small division,
// so the rates are lane-parallel over the compacted list.  Only the prefix
// chain is serial: one lane runs it from registers, 8 values per LDS block in
// two alternating register sets (the next block's loads in flight during the
// current block's adds).  Uniforms and the clock's -log(1 - u) are computed
// lane-parallel for 64 hops at a time.  The
// event search is monotone (the prefix never decreases and the quotient by
// the total is monotone), so two lane-parallel rounds find the first index
// (a coarse pass over 64 block ends, then the block).  With STAGE the layout's
// tables (1/d, tc, (I_0*R)/d) are copied to LDS once per trajectory.
#include "kmc_device.h"

namespace kmc {

// LDS carve-up for S sites, N acceptors (all offsets in doubles, 16-byte aligned).
struct PlainLds {
  size_t se, buf, lists, invd, tc, ird, total;
  __host__ __device__ PlainLds(int S, int N, bool stage) {
    auto up2 = [](size_t x) { return (x + 1) & ~(size_t)1; };
    const size_t SS = (size_t)S * S;
    se = 0;
    buf = up2(S);
    // possible pairs: nA (N - nA + P) + P (N - nA) <= N^2/4 + P N for nA occupied
    const size_t npairs = (size_t)N * N / 4 + (size_t)(S - N) * N + 16;
    lists = buf + up2(npairs);   // int16: occupied acceptors [N], targets [S], kept [npairs]
    invd = lists + up2(((size_t)N + S + npairs + 3) / 4);
    tc = invd + (stage ? up2((size_t)N * N) : 0);
    ird = tc + (stage ? up2(SS) : 0);
    total = ird + (stage ? up2(SS) : 0);
  }
};

// q = x / d and r = x - q*d for 0 <= x < 2^20, 1 <= d <= 256: an FP32 estimate
// from the reciprocal (relative error ~1e-7, so q is off by at most one), fixed.
__device__ __forceinline__ int div_small(int x, int d, float rcp, int& r) {
  int q = (int)((float)x * rcp);
  r = x - q * d;
  if (r < 0) {
    q -= 1;
    r += d;
  }
  if (r >= d) {
    q += 1;
    r -= d;
  }
  return q;
}

// PHASE: the phase-timing instantiation (shader-clock reads at the phase
// boundaries of every hop; the work is unchanged).
template <bool REC, bool STAGE, bool PHASE>
__global__ void __launch_bounds__(64) kmc_plain_fp64_kernel(PlainParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.x;
  const int lane = threadIdx.x;
  const TrajDev td = p.trajs[t];
  const LayoutDev L = p.layouts[td.layout];
  const EsetDev es = p.esets[td.eset];
  const int N = L.N, P = L.P, S = L.S, SS = S * S;
  const double nu = L.nu, kT = L.kT;
  const double nIR = (-L.I0) * L.R;
  // kT == 1 and nu == 1 make -dE/kT and nu*exp(x) exact copies of their
  // operand: the same bits with or without the operation
  const bool kT1 = kT == 1.0, nu1 = nu == 1.0;

  double* dl = reinterpret_cast<double*>(smem);
  const PlainLds off(S, N, STAGE);
  double* se = dl + off.se;
  double* buf = dl + off.buf;     // rates of the possible pairs, then their prefix sums
  int16_t* srcA = reinterpret_cast<int16_t*>(dl + off.lists);   // occupied acceptors
  int16_t* tgt = srcA + N;        // empty acceptors, then electrodes N..S-1
  int16_t* kidx = tgt + S;        // list positions of the kept rates
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
  // lanes of site chunk c (site lane + 64c) holding acceptors; S <= 80 < 128
  const uint64_t acc0 = N >= 64 ? ~0ull : ((1ull << N) - 1);
  const uint64_t acc1 = N <= 64 ? 0ull : ((1ull << (N - 64)) - 1);

  for (int e = lane; e < P; e += 64) se[N + e] = es.V[e];
  double E0 = 0.0, E1 = 0.0;
  if (lane < N) E0 = es.E[lane];
  if (lane + 64 < N) E1 = es.E[lane + 64];
  wave_lds_sync();
  int64_t eo = 0;                 // lane e < P holds electrode e's tally
  double ot0 = 0.0, ot1 = 0.0;    // post-hop occupation time, sites lane and lane+64
  double time = 0.0;
  int32_t status = ST_OK;
  const int64_t total_hops = td.prehops + td.hops;
  int64_t h = 0;
  uint64_t ph[PLAIN_PHASES] = {};
  uint64_t tp = PHASE ? __builtin_amdgcn_s_memtime() : 0;
  // cycles since the last mark into phase k
  auto mark = [&](int k) {
    if constexpr (PHASE) {
      const uint64_t now = __builtin_amdgcn_s_memtime();
      ph[k] += now - tp;
      tp = now;
    }
  };
  // -log(1 - u_clock) and u_event of hop (h & ~63) + lane: 64 hops of uniforms
  // and clock logarithms at a time, read with readlane per hop (the same
  // values and operations as per hop)
  double lg_lane = 0.0, ue_lane = 0.0;
  auto refill = [&](int64_t hh) {
    const int64_t hk = (hh & ~(int64_t)63) + lane;
    double uc = 0.5, ue = 0.5;
    if (p.uniforms) {
      if (hk < total_hops) {
        uc = p.uniforms[2 * hk];
        ue = p.uniforms[2 * hk + 1];
      }
    } else {
      stream_uniforms(td.seed, td.hop0 + (uint64_t)hk, uc, ue);
    }
    lg_lane = -log(1.0 - uc);
    ue_lane = ue;
  };
  // every scalar load of the set-up has returned before the loop: LDS
  // operations alone then complete in order, so the waits for the prefix
  // chain's blocks count only the operations issued after each block
  __builtin_amdgcn_s_waitcnt(0x0070);   // vmcnt(0) lgkmcnt(0)
  for (; h < total_hops; ++h) {
    const bool tally = h >= td.prehops;
    if (h == td.prehops) {   // reset() between prehops and hops (:562-585)
      time = 0.0;
      eo = 0;
      ot0 = ot1 = 0.0;
    }
    if ((h & 63) == 0) refill(h);
    // occupied / empty acceptors as wave-uniform masks of the two site chunks
    const uint64_t o0 = occ.a & acc0, o1 = occ.b & acc1;
    const uint64_t e0 = acc0 & ~o0, e1 = acc1 & ~o1;
    // the lists: occupied acceptors, then targets (empty acceptors, electrodes)
    const int nA = __builtin_popcountll(o0) + __builtin_popcountll(o1);
    const int cE = N - nA;
    const int cA = cE + P;
    {
      const bool oc0 = (o0 >> lane) & 1ull, oc1 = (o1 >> lane) & 1ull;
      const bool em0 = (e0 >> lane) & 1ull, em1 = (e1 >> lane) & 1ull;
      const int lo_o = __builtin_popcountll(o0), lo_e = __builtin_popcountll(e0);
      if (oc0) srcA[mbcnt(o0)] = (int16_t)lane;
      if (oc1) srcA[lo_o + mbcnt(o1)] = (int16_t)(lane + 64);
      if (em0) tgt[mbcnt(e0)] = (int16_t)lane;
      if (em1) tgt[lo_e + mbcnt(e1)] = (int16_t)(lane + 64);
      if (lane < P) tgt[cE + lane] = (int16_t)(N + lane);
    }
    const int nAA = nA * cA;
    const int n = nAA + P * cE;
    wave_lds_sync();
    mark(1);
    // site energies, :59-66: the empty acceptors j (tgt[0..cE), ascending),
    // four at a time with their loads issued together, added in order
    {
      double a0 = 0.0, a1 = 0.0;
      const int l0 = min(lane, N - 1), l1 = min(lane + 64, N - 1);
      for (int q0 = 0; q0 < cE; q0 += 4) {
        const int cnt = min(4, cE - q0);
        int j[4];
        double x0[4], x1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) j[q] = tgt[q0 + min(q, cnt - 1)];
#pragma unroll
        for (int q = 0; q < 4; ++q) x0[q] = invd[j[q] * N + l0];
        if (N > 64) {
#pragma unroll
          for (int q = 0; q < 4; ++q) x1[q] = invd[j[q] * N + l1];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < cnt) {
            a0 += x0[q];
            if (N > 64) a1 += x1[q];
          }
      }
      if (lane < N) se[lane] = E0 + nIR * a0;
      if (lane + 64 < N) se[lane + 64] = E1 + nIR * a1;
    }
    wave_lds_sync();
    mark(0);
    // rates of the possible pairs in row-major order, :69-87; two list entries
    // per lane and round, computed side by side (entries past n are evaluated
    // at n - 1 and dropped).
    //
    // Absorbed entries are dropped too.  All rates are >= 0, so the running
    // sum s_{k-1} before entry k is at least the largest earlier rate M, and a
    // rate below half an ulp of M (so below half an ulp of s_{k-1}) leaves the
    // rounded sum unchanged: fl(s + r) == s.  M is bounded from below by the
    // running maximum of the rates' high words (their low 32 bits cleared),
    // a wave max-scan on 32-bit integers; every partial sum the prefix chain
    // forms over the kept entries is the reference's partial sum at that
    // entry, and the dropped entries' partial sums equal the kept entry's
    // before them.  The kept rates go to buf[0..ns) and their list positions
    // to kidx.
    const float rA = cA > 0 ? __builtin_amdgcn_rcpf((float)cA) : 0.f;
    const float rE = cE > 0 ? __builtin_amdgcn_rcpf((float)cE) : 0.f;
    auto pair_of = [&](int idx, int& i, int& j) {
      int r;
      if (idx < nAA) {
        const int q = div_small(idx, cA, rA, r);
        i = srcA[q];
      } else {
        const int q = div_small(idx - nAA, cE, rE, r);
        i = N + q;
      }
      j = tgt[r];
    };
    auto rate_of = [&](int i, int j) {
      const int k = i * S + j;
      double dE;
      if (i < N && j < N) dE = se[j] - se[i] - irdp[k];
      else dE = se[j] - se[i];
      const double x = kT1 ? -dE : -dE / kT;
      const double ex = exp(x);
      const double m = dE > 0 ? (nu1 ? ex : nu * ex) : nu;
      return tcp[k] * m;
    };
    uint32_t hcarry = 0;   // running maximum of the high words (wave-uniform)
    int ns = 0;            // entries kept so far (wave-uniform)
    auto keep = [&](double r, int idx) {
      const bool valid = idx < n;
      const uint32_t hi = valid ? (uint32_t)((uint64_t)__double_as_longlong(r) >> 32) : 0u;
      const uint32_t m = max(wave_max_u32(hi), hcarry);
      hcarry = (uint32_t)__builtin_amdgcn_readlane((int)m, 63);
      const uint32_t e = (m >> 20) & 0x7FFu;   // biased exponent of the bound
      const uint32_t thr = e > 53u ? (e - 53u) << 20 : 0u;   // high word of its half ulp
      const bool kp = valid && hi >= thr;
      const uint64_t km = ballot(kp);
      if (kp) {
        const int pos = ns + mbcnt(km);
        buf[pos] = r;
        kidx[pos] = (int16_t)idx;
      }
      ns += __builtin_popcountll(km);
    };
    for (int b0 = 0; b0 < n; b0 += 128) {
      const int ia = b0 + lane, ib = b0 + 64 + lane;
      int i0, j0, i1, j1;
      pair_of(ia < n ? ia : n - 1, i0, j0);
      pair_of(ib < n ? ib : n - 1, i1, j1);
      const double r0 = rate_of(i0, j0);
      const double r1 = rate_of(i1, j1);
      keep(r0, ia);
      keep(r1, ib);
    }
    wave_lds_sync();
    mark(2);
    // cumulative sum, :91-97, over the kept entries: one lane; blocks of 8 in
    // two register sets that alternate, so the next block's loads are in
    // flight during this block's adds
    double acc = 0.0;
    if (lane == 0) {
      const int nb = ns >> 3;
      double A[8], B[8];
      // the running sum through one block of 8, written back as it goes: the
      // block's registers are only read, so they are free once it is done
      auto chain = [&](const double (&v)[8], int base) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          acc = v[q] + acc;
          buf[base + q] = acc;
        }
      };
      if (nb > 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) A[q] = buf[q];
      }
      int blk = 0;
#pragma unroll 1
      for (; blk + 1 < nb; blk += 2) {
        // straight-line body: the reload reads an existing block even past the
        // last pair (its values are then unused)
        const int nxt = min(blk + 2, nb - 1) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) B[q] = buf[(blk + 1) * 8 + q];
        __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the adds
        chain(A, blk * 8);
#pragma unroll
        for (int q = 0; q < 8; ++q) A[q] = buf[nxt + q];
        __builtin_amdgcn_sched_barrier(0);
        chain(B, (blk + 1) * 8);
      }
      if (blk < nb) chain(A, blk * 8);
      for (int idx = nb << 3; idx < ns; ++idx) {
        acc = buf[idx] + acc;
        buf[idx] = acc;
      }
    }
    const double total = readlane_f64(acc, 0);
    mark(3);
    if (p.totals && t == 0 && lane == 0) p.totals[h] = total;
    if (total == 0.0) {   // reference: ZeroDivisionError at 1/problist[-1]
      status = ST_ZERO_RATE;
      break;
    }
    const double lg = readlane_f64(lg_lane, (int)(h & 63));
    const double ue = readlane_f64(ue_lane, (int)(h & 63));
    const double scale = 1.0 / total;
    const double hop_time = lg * scale;
    wave_lds_sync();
    mark(4);
    // first k with problist[k]/total >= u, :103-110 (monotone in k): a kept
    // entry, since a dropped one repeats the partial sum of the kept entry
    // before it
    int from = 0, to = 0;   // u == 0 and no match: the reference's event 0, (0, 0)
    if (ue != 0.0) {
      const int stride = (ns + 63) >> 6;
      const int end = min(ns, (lane + 1) * stride) - 1;
      const bool hit = lane * stride < ns && buf[end] / total >= ue;
      const uint64_t mc = ballot(hit);
      if (mc) {
        const int base = __builtin_ctzll(mc) * stride;
        const int pos = base + lane;
        const uint64_t mf = ballot(lane < stride && pos < ns && buf[pos] / total >= ue);
        pair_of(kidx[base + __builtin_ctzll(mf)], from, to);   // the block's last entry hits
      }
      // the same on every lane: keep the event and the occupation it updates
      // on the scalar unit
      from = __builtin_amdgcn_readfirstlane(from);
      to = __builtin_amdgcn_readfirstlane(to);
    }
    const int k_ev = from * S + to;
    mark(5);
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
    wave_lds_sync();
    mark(6);
  }
  if (PHASE && p.phases && t == 0 && lane == 0)
    for (int k = 0; k < PLAIN_PHASES; ++k) p.phases[k] = ph[k];
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

template __global__ void kmc_plain_fp64_kernel<false, false, false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<true, false, false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<false, true, false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<true, true, false>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<false, true, true>(PlainParams);
template __global__ void kmc_plain_fp64_kernel<false, false, true>(PlainParams);

// LDS bytes the kernel needs for S sites, N acceptors.
size_t plain_lds_bytes(int S, int N, bool stage) {
  return PlainLds(S, N, stage).total * sizeof(double);
}

}  // namespace kmc
