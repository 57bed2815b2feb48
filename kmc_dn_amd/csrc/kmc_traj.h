// FP32-rate / FP64-accumulate hop loop of one trajectory (device code shared by
// the solo kernels, kmc_batch.hip, and the team kernels, kmc_team.hip).
//
// Same model as _simulate_discrete_record (kmc_dopant_networks.py:33-135) and
// Go's simulate (goSimulation/simulation.go:194-325), organised for CDNA4:
//
//  * Each lane owns sites s = lane + 64*c (c < C).  Site energies live in
//    registers as e~ = se * log2(e)/kT; electrodes hold V * log2(e)/kT.
//  * With lt = log2(nu * tc) and c = scaled Coulomb term I_0*R/d a rate is
//        2^(lt + min(0, e~_src - e~_dst + c)) = 2^min(lt, b + e~_src - e~_dst),
//    b = lt + c: one v_exp_f32 per pair and no FP32 underflow of the 10^-50
//    prefactors.  Where the pair terms come from is a policy:
//      TabB  {lt, b} pair table + c table in LDS, shared by the W waves of a
//            workgroup that run W trajectories of one layout (fewest VALU ops);
//      TabC  {lt, c} pair table in LDS (smaller; largest networks);
//      Pos   positions only (16 B per site in LDS per wave): d, lt and c are
//            computed per pair, ~2.5x the VALU work but no table, so waves of
//            different layouts fill a CU independently.
//  * Per hop, the smaller of {sources} (occupied acceptors + electrodes) and
//    {targets} (empty acceptors + electrodes) is walked in blocks of four
//    members (the last block exactly the members left) while the lanes cover
//    every site of the other role: row sums R_lane accumulate in registers with
//    no cross-lane traffic.
//  * Table policies leave the electrodes out of the walked set: with
//    w = nu * tc[s][e] the rates between acceptor s and all electrodes sum to
//        sum_e w * min(1, 2^e~_s * 2^-V~_e)   (s occupied, hops into them) or
//        sum_e w * min(1, 2^-e~_s * 2^V~_e)   (s empty, hops out of them),
//    one product (its clamp is the min) and one FMA per electrode and no exp2
//    per pair (elec_sums).  The per-trajectory factors 2^+-V~_e sit in LDS.  In
//    one-wave kernels a hop on which no lane of the other role has its energy
//    inside the electrodes' voltage window takes per-lane sums fixed for the
//    trajectory instead.
//  * Event pick: FP64 wave prefix sum of R (DPP), Philox event uniform * total,
//    ballot for the first lane whose prefix exceeds it, then one lane-parallel
//    pass over that lane's pairs (FP32 DPP scan) for the partner site.
//  * Hop: occupation bits flip in SGPRs; electrode tallies stay in the
//    electrode lanes' registers; the clock is accumulated in FP64.  A new
//    state's energies get O(N) Coulomb column updates (makeJump,
//    simulation.go:107-130): in one-chunk solo code the last FP64 recompute's
//    energies plus one column per acceptor whose occupation changed since, in
//    ascending site order (a function of the occupation alone); elsewhere the
//    previous state's plus the rank-1 update of the hop.
//  * At every multiple of `recompute_interval` hops the energies are rebuilt in
//    FP64 from E_constant and the I_0*R/d table (the Go reference never
//    re-synchronises).
//  * Uniforms come from one Philox block per lane every 64 hops (lane k holds
//    hop h0+k), so the RNG costs two readlanes per hop.
//  * A trajectory runs the absolute hop range [h_begin, h_end) of its budget
//    and, with `resume`, starts from the state its previous launch wrote.
//    Ranges that start at multiples of the recompute interval continue the
//    trajectory exactly as one long launch would (same energies, same stream).
//  * A team of T waves (one workgroup) can run one trajectory: wave w walks
//    the members u of the walked set with u = w (mod T), the partial row sums
//    meet in LDS (one barrier per hop) and are added in wave order, so every
//    wave continues with the same bits; the rest of the hop is replicated.
#pragma once
#include "kmc_device.h"

namespace kmc {

// LDS floats of a layout's electrode weights: P rounded up to 4 rows (padding
// rows zero) of 64C lane sites, so every load of a block of four electrodes has
// an immediate offset.
__host__ __device__ inline int elw_lds_floats(int P, int C) { return ((P + 3) & ~3) * 64 * C; }

// Copies the electrode weights (global rows of 192 sites) into LDS rows of 64C.
template <int C>
__device__ __forceinline__ void load_elw(float* dst, const LayoutDev& L, int tid, int nthreads) {
  const int n = elw_lds_floats(L.P, C);
  for (int k = tid; k < n; k += nthreads) dst[k] = L.elw[(k / (64 * C)) * 192 + k % (64 * C)];
}

template <int C>
__device__ __forceinline__ void recompute_energies(float (&se)[C], const uint64_t (&occ)[C],
                                                   const LayoutDev& L, const EsetDev& es,
                                                   int lane) {
  const int N = L.N, S = L.S;
  const double sc = (double)L.scale2;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    float v = 0.f;
    if (s < N) {
      // the sum over the empty acceptors in ascending order; the terms of eight
      // acceptors are loaded together (one memory round trip per eight; an
      // index past the last loads the last's term and adds nothing)
      const double e = es.E[s];
      const double* col = L.coul64 + s;
      double acc = 0.0;
      if constexpr (C == 1) {
        for (int j0 = 0; j0 < N; j0 += 8) {
          double t[8];
#pragma unroll
          for (int g = 0; g < 8; ++g) t[g] = col[min(j0 + g, N - 1) * N];
#pragma unroll
          for (int g = 0; g < 8; ++g) {
            const int j = j0 + g;
            acc += (j < N && !((occ[0] >> (j & 63)) & 1ull) && j != s) ? t[g] : 0.0;
          }
        }
      } else {
        for (int j = 0; j < N; ++j)
          if (!((occ[j >> 6] >> (j & 63)) & 1ull) && j != s) acc += col[j * N];
      }
      v = (float)((e - acc) * sc);
    } else if (s < S) {
      v = (float)(es.V[s - N] * sc);
    }
    se[c] = v;
  }
}

// Lowest set bit of a wave-uniform mask, -1 when it is empty (s_ff1 returns -1
// for 0, which __builtin_ctzll leaves undefined).
__device__ __forceinline__ int lowest_bit(uint64_t m) {
  int r;
  asm("s_ff1_i32_b64 %0, %1" : "=s"(r) : "s"(m));
  return r;
}

// Set bits of a wave-uniform mask as a 32-bit scalar (s_bcnt1 sets SCC).
__device__ __forceinline__ int popcount64(uint64_t m) {
  int r;
  asm("s_bcnt1_i32_b64 %0, %1" : "=s"(r) : "s"(m) : "scc");
  return r;
}

// m with bit (b mod 64) cleared.
__device__ __forceinline__ uint64_t clear_bit(uint64_t m, int b) {
  asm("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(b));
  return m;
}

// Lanes l of site chunk c (site l + 64c) with lo <= site < hi, as a lane mask.
__device__ __forceinline__ uint64_t chunk_mask(int lo, int hi, int c) {
  const int a = lo - 64 * c < 0 ? 0 : lo - 64 * c;
  const int b = hi - 64 * c > 64 ? 64 : hi - 64 * c;
  if (b <= a) return 0ull;
  const uint64_t below_b = b >= 64 ? ~0ull : ((1ull << b) - 1);
  return below_b & ~((1ull << a) - 1);
}

// min(a, b) for non-NaN operands as one v_min_f32.  fminf (and fmed3 with an
// infinite operand, which the compiler folds into it) first canonicalises any
// operand loaded from LDS with an extra v_max_f32 under IEEE mode; the pair
// terms are never NaN, so the plain instruction is exact here.
__device__ __forceinline__ float min_f32(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// ---------------------------------------------------------------------------
// Pair-term policies.  For a uniform member u of the walked set and the lane
// sites s = lane + 64c, pair(pre(u), c, diff) is log2 of the rate of the pair
// (src, dst) with e~_src - e~_dst = diff; coul(u, c) is c(s, u) for
// the energy update (u an acceptor).  prefetch(u) issues the loads that
// pair() needs, so four members can be in flight together.
// ---------------------------------------------------------------------------
template <int C, bool BT>
struct TabPolicy {
  static constexpr bool kElec = true;   // electrodes in closed form (elec_sums)
  // fast hops and leaps over several site chunks (table workgroups are few
  // per CU: their LDS, not their registers, bounds the waves per SIMD)
  static constexpr bool kFastChunks = true;
  // coul(N, c) is 0: column N of the c table lies in its zero padding, and column
  // N of the pair table is an electrode column (or, with no electrodes, the
  // padding), c = 0
  static constexpr bool kZeroColumnN = true;
  // [S*S + 64C]: entry (s, u) at u*S + s; the 64C padding entries (lt = -inf)
  // take the reads of the lanes past S in column S - 1
  const float2* tab;
  const float* ctab;   // BT: [N*S + 64C] c at u*S + s
  const float* elw;    // [4*ceil(P/4)][64C] electrode weights of the lane sites
  float* evc;          // this wave's [2][PMAX]: 2^V~_e, then 2^-V~_e
  int S, lane;
  // column strides in bytes (S * 8, S * 4), held in vector registers so that a
  // member's column address is one v_mad_u32_u24 of its scalar index
  int stride2, stride1;
  __device__ __forceinline__ static TabPolicy make(const float2* tab, const float* ctab,
                                                   const float* elw, float* evc, int S, int lane) {
    TabPolicy p{tab, ctab, elw, evc, S, lane, S * 8, S * 4};
    asm("" : "+v"(p.stride2), "+v"(p.stride1));
    return p;
  }
  // the lane's entry of column u: tab + lane + u*S
  __device__ __forceinline__ const float2* col(int u) const {
    return reinterpret_cast<const float2*>(reinterpret_cast<const char*>(tab + lane) +
                                           __umul24((unsigned)u, (unsigned)stride2));
  }
  struct Pre {
    float2 e[C];
  };
  // u in [0, S)
  __device__ __forceinline__ Pre prefetch(int u) const {
    Pre p;
    const float2* cl = col(u);
#pragma unroll
    for (int c = 0; c < C; ++c) p.e[c] = cl[64 * c];
    return p;
  }
  // energy difference e~_src - e~_dst = sg * d (sg = +-1, exact)
  __device__ __forceinline__ float pair(const Pre& p, int c, float d, float sg) const {
    if constexpr (BT) return min_f32(p.e[c].x, fmaf(sg, d, p.e[c].y));
    else return p.e[c].x + min_f32(0.f, fmaf(sg, d, p.e[c].y));
  }
  __device__ __forceinline__ float coul(int u, int c) const {
    if constexpr (BT)
      return reinterpret_cast<const float*>(reinterpret_cast<const char*>(ctab + lane) +
                                            __umul24((unsigned)u, (unsigned)stride1))[64 * c];
    else return col(u)[64 * c].y;
  }
};

template <int C>
struct PosPolicy {
  static constexpr bool kElec = false;   // electrodes are walked as members
  // no fast hops over several site chunks: positions-form batches are bound
  // by the waves per SIMD their registers allow (148 against 227 VGPRs at
  // three chunks); one-chunk trajectories take the one-chunk code
  static constexpr bool kFastChunks = false;
  static constexpr bool kZeroColumnN = false;
  const float* elw = nullptr;
  float* evc = nullptr;
  const float4* pos;   // [S] per wave: x, y, z, 1 for acceptors / 0 for electrodes
  float px[C], py[C], pz[C], kcl[C];   // the lane sites' coordinates and c scale
  uint64_t elm[C];     // lanes of chunk c that hold electrodes
  float l0, mkl;       // lt = l0 + mkl * d
  int lane, N, S;
  struct Pre {
    float4 q;
    bool el;           // wave-uniform: the member is an electrode
  };
  __device__ __forceinline__ void init(const float4* p, const LayoutDev& L, int ln) {
    pos = p;
    lane = ln;
    N = L.N;
    S = L.S;
    l0 = L.pos_l0;
    mkl = L.pos_mkl;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      const float4 q = s < L.S ? pos[s] : make_float4(1e6f, 1e6f, 1e6f, 0.f);
      px[c] = q.x;
      py[c] = q.y;
      pz[c] = q.z;
      kcl[c] = q.w * L.pos_kc;
      elm[c] = chunk_mask(L.N, L.S, c);
    }
  }
  // u in [0, S)
  __device__ __forceinline__ Pre prefetch(int u) const {
    Pre p;
    p.q = pos[u];
    p.el = u >= N;
    return p;
  }
  __device__ __forceinline__ float dist_rinv(const float4& q, int c, float& rinv) const {
    const float dx = px[c] - q.x, dy = py[c] - q.y, dz = pz[c] - q.z;
    // explicit FMAs: the same bits in every instantiation, whatever the
    // compiler would contract around it
    const float d2 = fmaxf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)), 1e-30f);
    rinv = __builtin_amdgcn_rsqf(d2);
    return d2 * rinv;
  }
  // energy difference e~_src - e~_dst = sg * dd (sg = +-1, exact)
  __device__ __forceinline__ float pair(const Pre& p, int c, float dd, float sg) const {
    float rinv;
    const float d = dist_rinv(p.q, c, rinv);
    const float lt = fmaf(mkl, d, l0);
    const float cc = rinv * kcl[c] * p.q.w;
    const float x = lt + min_f32(0.f, fmaf(sg, dd, cc));
    // electrode-electrode pairs (an electrode with itself included) have no
    // rate: a wave-uniform lane mask selects them
    return __builtin_amdgcn_inverse_ballot_w64(p.el ? elm[c] : 0ull) ? -INFINITY : x;
  }
  __device__ __forceinline__ float coul(int u, int c) const {
    const float4 q = pos[u];
    float rinv;
    (void)dist_rinv(q, c, rinv);
    return (u == lane + 64 * c) ? 0.f : rinv * kcl[c] * q.w;
  }
};

// T waves sharing one trajectory.  red: LDS [2][T][C][64] partial row sums
// (double-buffered by hop parity, so one barrier per hop suffices).
template <int T, int C>
struct Team {
  int wave = 0;        // rank in the team (wave-uniform)
  float* red = nullptr;
  // members u of the walked set this wave takes: u = wave (mod T)
  __device__ __forceinline__ uint64_t own() const {
    if constexpr (T == 1) return ~0ull;
    else if constexpr (T == 2) return 0x5555555555555555ull << wave;
    else return 0x1111111111111111ull << wave;
  }
  __device__ __forceinline__ bool writer() const { return T == 1 || wave == 0; }
  // Sum of the T partial row sums, added in wave order (same bits in every wave).
  __device__ __forceinline__ void reduce(float (&R)[C], int parity, int lane) const {
    if constexpr (T > 1) {
      float* mine = red + ((parity * T + wave) * C) * 64;
#pragma unroll
      for (int c = 0; c < C; ++c) mine[c * 64 + lane] = R[c];
      __syncthreads();
      const float* all = red + (parity * T) * C * 64;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        float acc = all[c * 64 + lane];
#pragma unroll
        for (int w = 1; w < T; ++w) acc += all[(w * C + c) * 64 + lane];
        R[c] = acc;
      }
    }
  }
};

// K members of the walked set (the lowest K bits of m, all present; K <= the
// number of bits left): their columns are loaded together, then their terms added.
template <int K, int C, class Pol>
__device__ __forceinline__ void row_block(const Pol& pol, int S, int cu, uint64_t& m,
                                          const float (&v)[C], const float (&se)[C], float sg,
                                          float (&R)[C]) {
  int b[K];
  typename Pol::Pre pre[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    b[k] = lowest_bit(m);
    m = clear_bit(m, b[k]);
  }
#pragma unroll
  for (int k = 0; k < K; ++k) pre[k] = pol.prefetch(64 * cu + b[k]);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float su = readlane_f32(se[cu], b[k]);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (c > 0 && 64 * c >= S) break;   // chunk past this layout's sites
      R[c] += fast_exp2(pol.pair(pre[k], c, v[c] - su, sg));
    }
  }
}

// Row sums over the uniform set U (bits of umask & own), lanes covering the
// other role.  sg = +1: U = targets, lanes = sources (e~_src - e~_dst = v - su);
// sg = -1: U = sources, lanes = targets.  One code path for both: the sign is a
// scalar operand of the pair term's FMA.  Members go in blocks of four; the last
// block takes exactly the members left (a scalar branch on their count), so no
// lane computes a padding term.
template <int C, class Pol>
__device__ __forceinline__ void row_sums(const Pol& pol, int S, const uint64_t (&umask)[C],
                                         uint64_t own, const float (&v)[C],
                                         const float (&se)[C], float sg, float (&R)[C]) {
#pragma unroll
  for (int cu = 0; cu < C; ++cu) {
    uint64_t m = umask[cu] & own;
    while (m) {
      const int n = popcount64(m);
      if (n >= 4) row_block<4>(pol, S, cu, m, v, se, sg, R);
      else if (n == 3) row_block<3>(pol, S, cu, m, v, se, sg, R);
      else if (n == 2) row_block<2>(pol, S, cu, m, v, se, sg, R);
      else row_block<1>(pol, S, cu, m, v, se, sg, R);
    }
  }
}

// Closed-form electrode terms of the row sums (table policies).  v: the lane
// sites' scaled energies; the sums of lanes outside the role are discarded by
// the caller, and electrode lanes have w = 0.  The exponent is clamped to 126
// (then x -+ V~ >= 26 and the min is 1 either way); |V~| <= 100 keeps 2^-+V~
// finite and nonzero, so the product is never NaN, and min(1, product) of a
// non-negative product is the multiply's own clamp to [0, 1].
// Electrodes go in blocks of four (all loads of a block issued together); rows
// past P have w = 0 and factor 1.  A team splits the blocks: wave w takes the
// blocks j = w (mod T).  The blocks are unrolled to the PMAX / 4 a layout can
// have, so every load takes an immediate offset from one base per hop; a
// wave-uniform branch skips the blocks past P.
template <int C, int T, class Pol>
__device__ __forceinline__ void elec_sums(const Pol& pol, int P, int S, bool udst, int wave,
                                          const float (&v)[C], float (&R)[C]) {
  float ex[C];
#pragma unroll
  for (int c = 0; c < C; ++c) ex[c] = fast_exp2(min_f32(udst ? v[c] : -v[c], 126.f));
  const float* k = pol.evc + (udst ? PMAX : 0) + 4 * wave;
  const float* w = pol.elw + pol.lane + 4 * wave * 64 * C;
  const int nb = (P + 3) >> 2;
  constexpr int kBlocks = (PMAX / 4 + T - 1) / T;
#pragma unroll
  for (int i = 0; i < kBlocks; ++i) {
    if (wave + T * i >= nb) break;
    const float4 k4 = *reinterpret_cast<const float4*>(k + 4 * T * i);
    const float kq[4] = {k4.x, k4.y, k4.z, k4.w};
    float wq[4][C];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < C; ++c) wq[q][c] = w[(4 * T * i + q) * 64 * C + 64 * c];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (c > 0 && 64 * c >= S) break;
        R[c] = fmaf(wq[q][c], __builtin_amdgcn_fmed3f(ex[c] * kq[q], 0.f, 1.f), R[c]);
      }
  }
}

// (float)x of lane l: converted on every lane, then one readlane (the compiler
// would otherwise read both halves of x and convert the scalar pair)
__device__ __forceinline__ float readlane_as_f32(double x, int l) {
  float f = (float)x;
  asm("" : "+v"(f));
  return readlane_f32(f, l);
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

template <int C>
__device__ __forceinline__ double pick_lane_f64(const double (&x)[C], int site) {
  const int cs = site >> 6, ls = site & 63;
  double r = readlane_f64(x[0], ls);
#pragma unroll
  for (int c = 1; c < C; ++c)
    if (cs == c) r = readlane_f64(x[c], ls);
  return r;
}

// One state of a trajectory's two-entry state cache: what a hop in that
// state needs of it, all in registers.
template <int C>
struct StateEntry {
  uint64_t key[C];   // occupation words
  uint64_t nzr[C];   // lanes whose row sum is nonzero (in VGPRs: used on rare hops)
  uint64_t pnz[C];   // lanes whose rate in row psel is nonzero
  double tot;        // total rate (the FP64 prefix at the last lane), in a VGPR
  float rtot;        // 1 / (float)tot (v_rcp_f32), the clock's rate factor, in a VGPR
  double inc[C];     // FP64 inclusive prefix of the row sums over lanes
  double exc[C];     // inc - row sum: the prefix before the lane
  float se[C];       // FP32 site energies of the state
  float pcum[C];     // FP32 prefix of row psel's rates over its partners, in site order
  int psel;          // lane of that row (-2: none yet; never a lane or the -1 of no lane)
  int udst;          // rows are sources, the walked set is the targets
  int val;
  // The hop from this state back into the other entry's state, as the leaps
  // test it (uniform values held in VGPRs):
  int plast;         // its partner lane, in row psel; -1: unknown since either entry changed,
                     // -2: taken once since (the bounds are formed when it is taken again)
  double lo, hi;     // row psel is the selected row iff lo <= target < hi
  double exs;        // exc of row psel
  float plo, phi;    // lane plast is the selected partner iff plo <= rr < phi
  int dlt[C];        // per lane site: its change of the tallies (+1 destination, -1 source)
};

// x in VGPRs: an inline-asm result with a VGPR operand counts as varying
// across lanes, so x and what is computed from it stay in VGPRs.  For
// wave-uniform values that only vector instructions (or rare paths) use,
// keeping them out of the SGPRs that the hop loop runs short of.
__device__ __forceinline__ double in_vgpr(double x) {
  asm("" : "+v"(x));
  return x;
}
__device__ __forceinline__ uint64_t in_vgpr(uint64_t x) {
  asm("" : "+v"(x));
  return x;
}
__device__ __forceinline__ int in_vgpr_i(int x) {
  asm("" : "+v"(x));
  return x;
}
__device__ __forceinline__ float in_vgpr(float x) {
  asm("" : "+v"(x));
  return x;
}
// A wave-uniform 64-bit value held in VGPRs, read into SGPRs.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// The hop loop keeps a state cache of two entries: the current state and the
// state before it.  Every hop changes the occupation, so the next state is
// either the one before (a hit: nothing is evaluated, the cached row sums,
// FP64 prefix, energies and, for the same selected row, the partner prefix
// serve the hop) or new (a miss: its energies are formed from the last
// recompute's, and its row sums are evaluated into the older entry).  The
// entries swap roles on every hop; the loop is unrolled by two hops, one for
// each assignment of the roles, so the entries sit in fixed registers and a
// hit moves no data.  The entries are emptied at every FP64 recompute, so
// ranges that start at multiples of the recompute interval continue exactly.
// The Go library keeps a cache keyed by occupation too
// (simulation.go:254-296); at low bias a trajectory mostly moves one hole
// back and forth between two sites, and the two entries serve ~99 % of its
// hops.  In a team every wave keeps the same cache (the reduced row sums are
// the same bits in every wave), so all waves hit together.
// PH: the phase-timing instantiation: shader-clock reads at the phase
// boundaries of every hop (FP32_PHASES), summed into p.phases by the first wave
// of each trajectory.  The reads wait for the wave's outstanding LDS and scalar
// loads, so its hops are somewhat slower than the others'.
//
// A one-chunk solo trajectory also keeps the states its two entries evict in a
// state store of p.store_slots slots in LDS (`store`, this wave's; null: none):
// a new state found there is swapped back into the entry instead of evaluated.
// The energies of a state are a function of its occupation alone between two
// recomputes (the recompute's energies plus the Coulomb columns of the
// acceptors whose occupation differs from the recompute's, added in ascending
// site order), so a stored entry holds the bits an evaluation would give: the
// store changes no result, only how often a state is evaluated.  Trajectories
// that hop among three or four states (at low bias, the slowest of an
// ensemble) miss the two entries on several per cent of their hops.
// ONE: the code runs in a kernel built for one site chunk (false for the
// one-chunk code of kernels over several chunks).
template <int C, bool REC, class Pol, int T = 1, bool PRIO = true, bool PH = false,
          bool ONE = (C == 1)>
__device__ void run_trajectory(const BatchParams& p, int t, const LayoutDev& L, const Pol& pol,
                               int lane, const Team<T, C>& team = Team<T, C>{},
                               unsigned char* store = nullptr) {

  const int64_t t_start = wall_clock64();
  const TrajDev td = p.trajs[t];
  const EsetDev es = p.esets[td.eset];
  const int N = L.N, S = L.S;
  TrajOut* o = p.out + t;
  // hop counters are 32-bit (prehops + hops <= INT32_MAX is checked when a batch
  // is created), so their comparisons stay on the scalar unit
  const int64_t total64 = td.prehops + td.hops;
  const int64_t hb64 = p.resume ? (td.h_begin < total64 ? td.h_begin : total64) : 0;
  const int64_t he64 = p.resume ? (td.h_end < total64 ? td.h_end : total64) : total64;
  const int h_begin = (int)hb64, h_end = (int)he64, prehops = (int)td.prehops;
  const bool cont = p.resume && h_begin > 0;

  uint64_t occ[C];   // acceptors (N <= 64C) live in the first C words
  float se[C];
  int eo[C];
  double ot[C];
#pragma unroll
  for (int w = 0; w < C; ++w) occ[w] = cont ? o->occ[w] : td.occ[w];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    eo[c] = (cont && s >= N && s < S) ? (int)o->eo[s - N] : 0;
    ot[c] = (cont && REC && p.occ_time && s < N) ? p.occ_time[(size_t)t * NMAX + s] : 0.0;
  }
  double time = cont ? o->time : 0.0;
  // The clock: each hop's FP32 waiting time waits in its hop's lane of the
  // uniform block (lane h & 63) until a fixed-order FP64 wave sum adds the block
  // to `time`, at every special hop (uniforms, end of the prehops, recompute)
  // and at the end of the range, so its bits do not depend on how the block's
  // hops were taken (general step, fast hop or leap).
  float dtv = 0.f;
  auto flush_clock = [&]() __attribute__((always_inline)) {
    time += readlane_f64(wave_scan_f64((double)dtv), 63);
    dtv = 0.f;
  };
  int32_t status = __builtin_amdgcn_readfirstlane(cont ? o->status : (int32_t)ST_OK);
  recompute_energies<C>(se, occ, L, es, lane);
  // the last recompute's energies and occupation: every state's energies derive
  // from them
  float base_se[C];
  uint64_t base_occ[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    base_se[c] = se[c];
    base_occ[c] = occ[c];
  }
  uint64_t accm[C], elm[C];   // lanes of chunk c holding acceptors / electrodes
#pragma unroll
  for (int c = 0; c < C; ++c) {
    accm[c] = chunk_mask(0, N, c);
    elm[c] = chunk_mask(N, S, c);
  }
  // electrodes in closed form when the policy has the weights and every scaled
  // voltage satisfies |V~| <= 100; otherwise they are walked as members
  bool elec = false;
  // Hops on which every lane of the other role has all its electrode terms
  // saturated (x >= max V~ side) or all below 1 take per-lane sums fixed for the
  // trajectory: sum_e w, or 2^x sum_e w 2^-+V~ (x = +-v as in elec_sums).
  float esum_w[C], esum_dst[C], esum_src[C];   // sum w, sum w 2^-V~, sum w 2^V~
  float vt_max = 0.f, vt_min = 0.f;
  if constexpr (Pol::kElec) {
    const float vt = lane < L.P ? (float)(es.V[lane] * (double)L.scale2) : 0.f;
    elec = L.P > 0 && L.elw_ok && ballot(!(fabsf(vt) <= 100.f)) == 0;
    if (lane < PMAX) {   // 2^0 = 1 past P: finite factors for the zero padding rows
      pol.evc[lane] = fast_exp2(vt);
      pol.evc[PMAX + lane] = fast_exp2(-vt);
    }
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < C; ++c) esum_w[c] = esum_dst[c] = esum_src[c] = 0.f;
    if (elec && T == 1) {
      vt_max = -INFINITY;
      vt_min = INFINITY;
      for (int e = 0; e < L.P; ++e) {
        const float ve = readlane_f32(vt, e);
        vt_max = fmaxf(vt_max, ve);
        vt_min = fminf(vt_min, ve);
        const float ks = pol.evc[e], kd = pol.evc[PMAX + e];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const float w = pol.elw[e * 64 * C + 64 * c + lane];
          esum_w[c] += w;
          esum_dst[c] = fmaf(w, kd, esum_dst[c]);
          esum_src[c] = fmaf(w, ks, esum_src[c]);
        }
      }
    }
  }
  uint64_t walk[C];   // sites that may be walked: acceptors, plus electrodes unless elec
#pragma unroll
  for (int c = 0; c < C; ++c) walk[c] = elec ? accm[c] : ~0ull;

  const int K = p.recompute_interval;
  float lg_lane = 0.f;    // log2(1 - u_clock) * -ln 2 = -ln(1 - u_clock) of hop (h & ~63) + lane
  double ue_lane = 0.0;   // event uniform of that hop
  // next FP64 recompute (a multiple of K; INT32_MAX: none left in this range)
  auto next_multiple = [K](int64_t x) {
    const int64_t n = (x / K + 1) * K;
    return (int)(n < INT32_MAX ? n : INT32_MAX);
  };
  int next_rc = next_multiple(h_begin);
  int h_stop = status == ST_OK ? h_end : h_begin;   // a stopped trajectory stays
  int h = h_begin;
  // uniforms of hops (hh & ~63) + lane: at the start and at every multiple of 64
  auto refill = [&](int hh) {
    double uc;
    stream_uniforms(td.seed, td.hop0 + (uint64_t)((hh & ~(int64_t)63) + lane), uc, ue_lane);
    lg_lane = fast_log2((float)(1.0 - uc)) * -0.69314718055994531f;
  };
  if (h < h_stop && (h & 63) != 0) refill(h);
  // the next hop with work besides the hop itself: a multiple of 64 (uniforms),
  // the end of the prehops (tallies restart) or a multiple of K (recompute)
  // (32-bit unsigned arithmetic: hh <= INT32_MAX, so r <= 2^31 and no 64-bit
  // compare, which the scalar unit lacks, is needed)
  auto next_special = [&](int hh, bool inclusive) {
    const unsigned r = inclusive ? ((unsigned)hh + 63u) & ~63u : ((unsigned)hh | 63u) + 1u;
    int n = (int)(r < (unsigned)INT32_MAX ? r : (unsigned)INT32_MAX);
    if (next_rc < n) n = next_rc;
    if ((inclusive ? prehops >= hh : prehops > hh) && prehops < n) n = prehops;
    return n;
  };
  int special = next_special(h, true);
  // hops below h_lim need nothing besides the hop itself (the fast path's range)
  int h_lim = special < h_stop ? special : h_stop;
  double b15 = 0.0, b31 = 0.0;   // row-broadcast registers of the FP64 scan
  uint64_t ph[FP32_PHASES] = {};
  uint64_t tp = PH ? __builtin_amdgcn_s_memtime() : 0;
  auto mark = [&](int k) {
    if constexpr (PH) {
      const uint64_t now = __builtin_amdgcn_s_memtime();
      ph[k] += now - tp;
      tp = now;
    }
  };
  int nred = 0;   // team reductions so far
  // Fast hops and leaps: kernels that record take the general step on every
  // hop, and so do teams over several site chunks, whose miss-heavy long
  // trajectories ran their general steps a fifth slower beside the fast path's
  // registers, and the positions form over several chunks
  constexpr bool kFast = !REC && (C == 1 || (T == 1 && Pol::kFastChunks));
  StateEntry<C> E0, E1;
  E0.val = E1.val = 0;
  E0.plast = E1.plast = -1;
  bool cur_ok = false;   // the current state's entry holds it
  // New states of the current recompute interval, and whether the last interval
  // had more than 16 (held in VGPRs: touched on new states and at recomputes
  // only, so they take no scalar registers from the hop loop).  Such a
  // trajectory hops among three or four states and is among the slowest of a
  // batch, whose launch lasts as long as they do: its waves issue at priority
  // 2 instead of 1 outside the hop's chain, ahead of the faster waves that
  // share their SIMDs.  One-chunk kernels with graded priority only: in an
  // uneven batch the solo kernel's miss-heavy waves would go ahead of the team
  // waves that carry its longest trajectories (3.5 % longer).  No result
  // depends on it.
  constexpr bool kSlowPrio = PRIO && ONE && T == 1;
  int nnew = in_vgpr_i(0);
  int slowv = in_vgpr_i(0);
  // Energies of a new state.  One-chunk solo code: the recompute's plus the
  // columns of the acceptors changed since, a function of the occupation alone
  // (what lets the state store restore a state's bits).  Teams and kernels
  // over several site chunks: the previous state's plus the rank-1 update of
  // the hop that left it (two columns; the miss-heavy trajectories these run
  // change many sites between recomputes); psrc, pdst: that hop's ends, pend:
  // its update is still to be applied.
  constexpr bool kPathFree = C == 1 && T == 1;
  int psrc = 0, pdst = 0;
  bool pend = false;
  // the state store: slots in use (bit v: slot v) and the slot the next evicted
  // state without a slot of its own takes (round robin)
  constexpr bool kStore = kPathFree;
  const int nslots = kStore && store ? p.store_slots : 0;
  unsigned svalid = 0;
  int snext = 0;
  uint64_t* const skey = reinterpret_cast<uint64_t*>(store);
  uint64_t* const snzr = skey + STORE_MAX_SLOTS;
  auto slot_inc = [&](int v) {
    return reinterpret_cast<double*>(store + STORE_HEADER_BYTES + v * STORE_SLOT_BYTES);
  };
  // entry W's state into slot v (W valid)
  auto store_put = [&](int v, const StateEntry<C>& W) __attribute__((always_inline)) {
    double* inc = slot_inc(v);
    inc[lane] = W.inc[0];
    inc[64 + lane] = W.exc[0];
    inc[128 + lane] = (double)W.se[0];   // (exact both ways)
    if (lane == 0) {
      skey[v] = W.key[0];
      snzr[v] = W.nzr[0];
    }
  };

  // The current state's row sums, FP64 prefix and total into W, from the
  // energies se; false: no rate is positive and finite.
  auto evaluate = [&](StateEntry<C>& W) __attribute__((always_inline)) -> bool {
    // roles of the lanes' sites, as wave-uniform lane masks (scalar ops only):
    // sources = occupied acceptors + electrodes, targets = empty acceptors + electrodes
    uint64_t srcm[C], dstm[C];
    int nocc = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const uint64_t o = occ[c] & accm[c];
      srcm[c] = o | elm[c];
      dstm[c] = (accm[c] ^ o) | elm[c];
      nocc += popcount64(o);
    }
    const bool udst = N - nocc <= nocc;   // walk the smaller set (empty vs occupied)
    // every lane sums over the walked members; the sums of lanes outside the
    // other role (and past S) are discarded after the reduction
    float R[C];
    uint64_t um[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      R[c] = 0.f;
      um[c] = (udst ? dstm[c] : srcm[c]) & walk[c];
    }
    const float (&v)[C] = se;
    if constexpr (Pol::kElec) {
      if (elec && T > 1) {
        // teams run large networks, where some lane is nearly always in the
        // window: no test
        elec_sums<C, T>(pol, L.P, S, udst, team.wave, v, R);
      } else if (elec) {
        // the lanes whose electrode terms are neither all saturated nor all
        // below 1 (acceptors of the other role); none: the per-lane sums
        const float thi = udst ? vt_max : -vt_min, tlo = udst ? vt_min : -vt_max;
        float x[C];
        uint64_t hi[C], mid = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          x[c] = min_f32(udst ? v[c] : -v[c], 126.f);
          hi[c] = ballot(x[c] >= thi);
          const uint64_t lo = ballot(x[c] < tlo);
          mid |= (udst ? srcm[c] : dstm[c]) & accm[c] & ~(hi[c] | lo);
        }
        if (mid == 0) {
#pragma unroll
          for (int c = 0; c < C; ++c)
            R[c] = __builtin_amdgcn_inverse_ballot_w64(hi[c])
                       ? esum_w[c] : fast_exp2(x[c]) * (udst ? esum_dst[c] : esum_src[c]);
        } else {
          elec_sums<C, T>(pol, L.P, S, udst, team.wave, v, R);
        }
      }
    }
    // +1: walked targets, e~_src - e~_dst = v - su; -1: walked sources
    const float sgr = __int_as_float(udst ? (int)0x3f800000 : (int)0xbf800000);
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
    row_sums<C>(pol, S, um, team.own(), v, se, sgr, R);
    // the team's partial sums alternate between two LDS buffers by reduction
    // (hops served from the state cache reduce nothing and pass no barrier;
    // the waves agree on every hit, so they count the same reductions)
    team.reduce(R, nred & 1, lane);
    ++nred;
#pragma unroll
    for (int c = 0; c < C; ++c)
      R[c] = __builtin_amdgcn_inverse_ballot_w64(udst ? srcm[c] : dstm[c]) ? R[c] : 0.f;
    mark(1);

    // From here through the next hop's electrode terms the work is one
    // dependency chain (scans, readlanes, the partner row's LDS loads unless
    // the state's entry holds its prefix, and a miss's Coulomb columns); the
    // waves sharing a SIMD fill its gaps with their row sums of the walked
    // members.  A raised issue priority lets the chain's instructions go as
    // soon as they are ready: 3 from the scan to the partner, 1 for the update
    // and the next hop's masks and electrode terms, 0 for the row sums (4096
    // trajectories: 12 % more hops per second).  PRIO false: the variant for
    // launches whose groups queue behind full SIMDs.
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(3);
    // FP64 prefix over lane rows
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (c > 0 && 64 * c >= S) {
        W.inc[c] = W.exc[c] = total;
        W.nzr[c] = in_vgpr((uint64_t)0);
        continue;
      }
      W.inc[c] = c == 0 ? wave_scan_f64((double)R[c])
                        : wave_scan_f64((double)R[c]) + total;
      W.exc[c] = W.inc[c] - (double)R[c];
      W.nzr[c] = in_vgpr(ballot(R[c] > 0.f));
      total = readlane_f64(W.inc[c], 63);
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      W.key[c] = occ[c];
      W.se[c] = se[c];
    }
    W.tot = in_vgpr(total);
    W.rtot = __builtin_amdgcn_rcpf((float)total);
    W.udst = udst;
    W.psel = -2;
    W.plast = -1;
    W.val = 1;
    mark(2);
    return total > 0.0 && total < INFINITY;
  };

  // The partner prefix of row `sel` of F's state into F: the prefix of the
  // row's rates over the lanes in site order, and the lanes whose rate is
  // nonzero (a lane-parallel pass over the row).
  auto partner_pass = [&](StateEntry<C>& F, int sel) __attribute__((always_inline)) {
    uint64_t inu[C];   // partners: targets of a source row, sources of a target row
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const uint64_t o = occ[c] & accm[c];
      inu[c] = (F.udst ? accm[c] ^ o : o) | elm[c];
    }
    const float s_sel = pick_lane_value<C>(F.se, sel);
    // sign of the lane-parallel difference: a scalar constant
    const float sg = __int_as_float(F.udst ? (int)0xbf800000 : (int)0x3f800000);
    const typename Pol::Pre pre = pol.prefetch(sel);
    float run = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      F.pcum[c] = 0.f;
      F.pnz[c] = 0;
      if (c > 0 && 64 * c >= S) continue;
      // udst ? s_sel - se : se - s_sel (the sign is exact)
      float r = fast_exp2(pol.pair(pre, c, F.se[c] - s_sel, sg));
      r = __builtin_amdgcn_inverse_ballot_w64(inu[c]) ? r : 0.f;
      F.pcum[c] = wave_scan_f32(r) + run;
      F.pnz[c] = ballot(r > 0.f);
      run = readlane_f32(F.pcum[c], 63);
    }
    F.psel = sel;
    F.plast = -1;
  };

  // One hop from the current state F (its entry valid); B: the other entry.
  // Returns 1 when the state it leads to is B's, 0 when that state is new,
  // -1 when the trajectory stops.
  auto hop = [&](StateEntry<C>& F, const StateEntry<C>& B) __attribute__((always_inline)) -> int {
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(3);
    // exponential waiting time -ln(1-u)/total, an FP32 quotient (the clock
    // keeps it in the hop's lane; the records take it in FP64)
    const float dt32 = lg_lane * F.rtot;
    const double dt = REC ? (double)readlane_f32(dt32, (int)(h & 63)) : 0.0;
    // event uniform x total, multiplied on every lane and read from this hop's
    const double target = readlane_f64(ue_lane * F.tot, (int)(h & 63));

    // the selected lane l and the remainder rr = target - (incl - R) at l, formed
    // on every lane (FP64 difference, rounded to FP32) and read from l
    int sel = -1;
    float rr = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const uint64_t m = ballot(F.inc[c] > target);
      if (sel < 0 && m) {
        const int l = __builtin_ctzll(m);
        sel = 64 * c + l;
        rr = readlane_as_f32(target - F.exc[c], l);
      }
    }
    if (sel < 0) {   // target rounded onto the total: last lane with a nonzero row
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const uint64_t m = uniform_u64(F.nzr[c]);
        if (m) {
          const int l = 63 - __builtin_clzll(m);
          sel = 64 * c + l;
          rr = readlane_as_f32(target - F.exc[c], l);
        }
      }
    }
    mark(3);

    // partner of `sel`: the prefix of its row over the lanes in site order (and
    // the lanes whose rate is nonzero), from the entry when it was formed for
    // the same row, else by a lane-parallel pass over the row
    if (sel != F.psel) {
      partner_pass(F, sel);
    } else if constexpr (PH) {
      ++ph[8];
    }
    int pick = -1, last_nz = -1;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (c > 0 && 64 * c >= S) break;
      const uint64_t hit = F.pnz[c] & ballot(F.pcum[c] > rr);
      if (pick < 0) {
        if (F.pnz[c]) last_nz = 64 * c + 63 - __builtin_clzll(F.pnz[c]);
        if (hit) pick = 64 * c + __builtin_ctzll(hit);
      }
    }
    if (pick < 0) pick = last_nz;
    if (pick < 0) return -1;
    mark(4);
    if constexpr (kSlowPrio) {
      if (__builtin_amdgcn_readfirstlane(slowv)) __builtin_amdgcn_s_setprio(2);
      else __builtin_amdgcn_s_setprio(1);
    } else if constexpr (PRIO) {
      __builtin_amdgcn_s_setprio(1);
    }
    const int src = F.udst ? sel : pick;
    const int dst = F.udst ? pick : sel;
    const bool tally = h >= prehops;

    if (REC && tally && !p.post_hop_record) {
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (__builtin_amdgcn_inverse_ballot_w64(occ[c] & accm[c])) ot[c] += dt;
    }
    // net traffic, Go convention (simulation.go:310-311)
    if (REC && p.traffic && tally && lane == 0 && team.writer()) {
      double* tr = p.traffic + (size_t)t * p.traffic_stride;
      tr[src * S + dst] += 1.0;
      tr[dst * S + src] -= 1.0;
    }
    // perform hop: occupation (acceptors live in the words of the C chunks),
    // tallies; a new state's energies are formed when it is evaluated
    const bool src_acc = src < N, dst_acc = dst < N;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      // the source's bit is cleared whether it is an acceptor or not: bits of
      // electrode sites are never set (a one-chunk kernel has S <= 64 sites)
      if (C == 1 || (src >> 6) == c) occ[c] &= ~(1ull << (src & 63));
      if (dst_acc && (C == 1 || (dst >> 6) == c)) occ[c] |= 1ull << (dst & 63);
    }
    psrc = src;
    pdst = dst;
    // electrode tallies (a wave-uniform branch: most hops join two acceptors);
    // they move on the lanes of both ends; only electrode lanes' are kept
    if (__builtin_expect(!src_acc || !dst_acc, 0)) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (lane + 64 * c == src) eo[c] -= 1;
        if (lane + 64 * c == dst) eo[c] += 1;
      }
    }
    if (REC && tally && p.post_hop_record) {
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (__builtin_amdgcn_inverse_ballot_w64(occ[c] & accm[c])) ot[c] += dt;
    }
    if (lane == (h & 63)) dtv = dt32;
    mark(5);
    bool same = B.val != 0;
#pragma unroll
    for (int c = 0; c < C; ++c) same = same && B.key[c] == occ[c];
    return same ? 1 : 0;
  };

  // The work of a special hop before the hop itself: the clock's block sum,
  // uniforms, the end of the prehops, the FP64 recompute (which empties the
  // state cache).
  // false: the launch's time limit has passed at this recompute point, and the
  // trajectory stops here (a multiple of the recompute interval, from which a
  // resumed range continues exactly as one long launch would).  Team waves read
  // their clocks apart and could disagree, so teams carry no limit (the host
  // refuses one).
  auto special_work = [&]() __attribute__((always_inline)) -> bool {
    if (T == 1 && p.time_limit > 0 && h == next_rc && h > h_begin &&
        wall_clock64() - t_start > p.time_limit) {
      h_stop = h;
      h_lim = h;
      return false;
    }
    flush_clock();
    if ((h & 63) == 0) refill(h);
    if (h == prehops) {
      time = 0.0;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        eo[c] = 0;
        ot[c] = 0.0;
      }
    }
    if (h == next_rc) {
      recompute_energies<C>(se, occ, L, es, lane);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        base_se[c] = se[c];
        base_occ[c] = occ[c];
      }
      next_rc = next_multiple(h);
      pend = false;
      cur_ok = false;
      E0.val = E1.val = 0;
      svalid = 0;
      snext = 0;
      if constexpr (kSlowPrio) {
        slowv = in_vgpr_i(nnew > 16 ? 1 : 0);
        nnew = in_vgpr_i(0);
        if (__builtin_amdgcn_readfirstlane(slowv)) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(1);
      }
    }
    special = next_special(h, false);
    h_lim = special < h_stop ? special : h_stop;
    mark(0);
    return true;
  };

  // One hop whose current state's entry is F (B: the entry of the state
  // before), its special work done; false: the trajectory stops here or its
  // range ends.
  auto step = [&](StateEntry<C>& F, StateEntry<C>& B) __attribute__((always_inline)) -> bool {
    if constexpr (!kFast) {
      if (__builtin_expect(h == special, 0) && !special_work()) return false;
    }
    if (!cur_ok) {
      if constexpr (kSlowPrio) nnew = in_vgpr_i(nnew + 1);
      // a new state (F's old state, two hops back, leaves the entries)
      B.plast = -1;   // B's hop back led into the state that F held
      bool restored = false;
      if constexpr (kStore) {
        if (nslots > 0) {
          mark(1);   // the time since the last hop: a miss's, whether restored or evaluated
          // the slots' keys by broadcast reads (a uniform address: no lane
          // address to keep in a register through the hop loop)
          unsigned found = 0;
#pragma unroll
          for (int v = 0; v < STORE_MAX_SLOTS; ++v)
            found |= (uniform_u64(skey[v]) == occ[0] ? 1u : 0u) << v;
          found &= svalid;
          if (found) {
            // swap: the stored state into F, F's old state into its slot
            const int v = __builtin_ctz(found);
            double* inc = slot_inc(v);
            const double si = inc[lane], sx = inc[64 + lane];
            const float sv = (float)inc[128 + lane];
            const uint64_t nz = snzr[v];
            if (F.val) store_put(v, F);
            else svalid &= ~(1u << v);
            int nocc = popcount64(occ[0] & accm[0]);
            F.inc[0] = si;
            F.exc[0] = sx;
            F.se[0] = sv;
            F.nzr[0] = in_vgpr(nz);
            F.tot = in_vgpr(readlane_f64(si, 63));
            F.rtot = __builtin_amdgcn_rcpf((float)F.tot);
            F.key[0] = occ[0];
            F.udst = N - nocc <= nocc;
            F.psel = -2;
            F.plast = -1;
            F.val = 1;
            restored = true;
            if constexpr (PH) ++ph[10];
            mark(2);
          } else if (F.val) {
            store_put(snext, F);
            svalid |= 1u << snext;
            snext = snext + 1 == nslots ? 0 : snext + 1;
          }
        }
      }
      if (!restored) {
        if constexpr (kPathFree) {
        // its energies: the recompute's plus the Coulomb columns of the
        // acceptors whose occupation differs from the recompute's (+ where
        // now occupied, - where now empty), added in ascending site order;
        // the columns of up to G sites are loaded together (a site past the
        // last loads the first's column and adds nothing)
        constexpr int G = 8;
#pragma unroll
        for (int c = 0; c < C; ++c) se[c] = base_se[c];
#pragma unroll
        for (int w = 0; w < C; ++w) {
          uint64_t dm = (occ[w] ^ base_occ[w]) & accm[w];
          while (dm) {
            int js[G];
            const int j0 = __builtin_ctzll(dm);
#pragma unroll
            for (int g = 0; g < G; ++g) {
              js[g] = dm ? __builtin_ctzll(dm) : -1;
              dm &= dm - 1;
            }
            float cv[G][C];
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
              for (int c = 0; c < C; ++c)
                cv[g][c] = (c == 0 || 64 * c < S) ? pol.coul(64 * w + (js[g] < 0 ? j0 : js[g]), c)
                                                  : 0.f;
#pragma unroll
            for (int g = 0; g < G; ++g) {
              if (js[g] < 0) break;
              const bool on = (occ[w] >> js[g]) & 1ull;
#pragma unroll
              for (int c = 0; c < C; ++c) se[c] += on ? cv[g][c] : -cv[g][c];
            }
          }
        }
        } else if (pend) {
          // the previous state's energies plus the last hop's rank-1 update
          // (both Coulomb columns loaded together; an electrode end
          // contributes 0); after a recompute, the recomputed ones
          const bool src_acc = psrc < N, dst_acc = pdst < N;
          float cs[C], cd[C];
#pragma unroll
          for (int c = 0; c < C; ++c) {
            if constexpr (Pol::kZeroColumnN) {   // column N holds c = 0: no select per lane
              cs[c] = (c == 0 || 64 * c < S) ? pol.coul(min((unsigned)psrc, (unsigned)N), c) : 0.f;
              cd[c] = (c == 0 || 64 * c < S) ? pol.coul(min((unsigned)pdst, (unsigned)N), c) : 0.f;
            } else {
              cs[c] = (c == 0 || 64 * c < S) && src_acc ? pol.coul(psrc, c) : 0.f;
              cd[c] = (c == 0 || 64 * c < S) && dst_acc ? pol.coul(pdst, c) : 0.f;
            }
          }
#pragma unroll
          for (int c = 0; c < C; ++c) se[c] = (B.se[c] - cs[c]) + cd[c];
        }
        if (!evaluate(F)) {
          status = ST_ZERO_RATE;
          return false;
        }
      }
      cur_ok = true;
    } else if constexpr (PH) {
      ++ph[7];
    }
    const int r = hop(F, B);
    if (r < 0) {
      status = ST_ZERO_RATE;
      return false;
    }
    cur_ok = r > 0;
    pend = true;
    ++h;
    mark(5);
    return h < h_stop;
  };
  // A hit hop in its shortest form: the current state's entry F is valid, its
  // partner prefix was formed for the row the event falls in and the hop leads
  // back into the state of the other entry B.  The arithmetic is the general
  // hop's (same target, remainder, partner and clock bits); only the
  // bookkeeping that such a hop cannot need is left out, so the scalar unit
  // that the waves of a CU share issues a third of the general path's
  // instructions.  Anything else (an uncached row, a new state, the end of the
  // range, uniforms or a recompute due) returns false with nothing changed, and
  // the general step takes the hop.  (Taking the hops into an uncached row or a
  // new state here too made the common path's code longer: 7 % slower.)
  // Kernels that record take the general step on every hop.
  //
  // Leaps: once the hop back is known from both states
  // (plast), whether hop k of the current uniform block takes it depends only
  // on its own uniform, so the lanes test the hops of the block in parallel,
  // lane k for hop (h & ~63) + k in the state that hop starts from (F and B
  // alternate): target in [lo, hi) selects the row, rr in [plo, phi) the
  // partner, with the general hop's operations.  The run of consecutive hops
  // back, cut to an odd length (the state then is B's), is applied at once: the
  // clock adds the hops' waiting times one by one in hop order, the tallies
  // change by the count of hops each way, the occupation is B's.  The bits are
  // those of taking the hops one at a time.
  auto fast_hop = [&](StateEntry<C>& F, const StateEntry<C>& B)
                      __attribute__((always_inline)) -> bool {
    if (h >= h_lim) return false;
    const int hl = h & 63;
    const double tlane = ue_lane * F.tot;   // every lane's hop's target in state F
    {
      if ((F.plast | B.plast) >= 0) {
        const double tb = ue_lane * B.tot;
        const float rf = (float)(tlane - F.exs), rb = (float)(tb - B.exs);
        const uint64_t okf =
            ballot(tlane >= F.lo && tlane < F.hi && rf >= F.plo && rf < F.phi);
        const uint64_t okb = ballot(tb >= B.lo && tb < B.hi && rb >= B.plo && rb < B.phi);
        // lanes k with k - hl even start their hop in state F
        const uint64_t inf = (hl & 1) ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
        const uint64_t run = ((okf & inf) | (okb & ~inf)) >> hl;
        // trailing ones (s_ff1 of an all-ones complement is -1: the whole block)
        int n = lowest_bit(~run);
        const int nmax = 64 - hl < h_lim - h ? 64 - hl : h_lim - h;
        if ((unsigned)n > (unsigned)nmax) n = nmax;
        if (n == 0) return false;
        // the run applied in a block of its own (n > 0 always holds here): the
        // register allocator then spills less in the hop loop (60 against 92 B)
        if (n > 0) {
        n -= (n & 1) ^ 1;
        const uint64_t lanes = (~0ull << hl) & (hl + n < 64 ? ~(~0ull << (hl + n)) : ~0ull);
        const float dtl = lg_lane * (__builtin_amdgcn_inverse_ballot_w64(inf) ? F.rtot : B.rtot);
        if (__builtin_amdgcn_inverse_ballot_w64(lanes)) dtv = dtl;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          eo[c] += ((n + 1) >> 1) * F.dlt[c] + (n >> 1) * B.dlt[c];
          occ[c] = B.key[c];
        }
        h += n;
        if constexpr (PH) {
          ph[7] += n;
          ph[8] += n;
          ph[9] += n;
        }
        mark(6);
        return true;
        }
      }
    }
    const double target = readlane_f64(tlane, hl);
    // sel and pick are -1 when no lane qualifies (s_ff1 of an empty mask); a
    // readlane from lane -1 reads lane 63 and its value goes unused: psel is
    // never -1
    int sel = -1;
    float rr = 0.f;
    if constexpr (C == 1) {
      sel = lowest_bit(ballot(F.inc[0] > target));
      rr = readlane_as_f32(target - F.exc[0], sel);
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (c > 0 && 64 * c >= S) break;
        const uint64_t m = ballot(F.inc[c] > target);
        if (sel < 0 && m) {
          const int l = __builtin_ctzll(m);
          sel = 64 * c + l;
          rr = readlane_as_f32(target - F.exc[c], l);
        }
      }
    }
    if (sel < 0) return false;   // target rounded onto the total: the general step's fallback
    if (sel != F.psel) return false;
    int pick = -1;
    if constexpr (C == 1) {
      pick = lowest_bit(F.pnz[0] & ballot(F.pcum[0] > rr));
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (c > 0 && 64 * c >= S) break;
        const uint64_t hit = F.pnz[c] & ballot(F.pcum[c] > rr);
        if (pick < 0 && hit) pick = 64 * c + __builtin_ctzll(hit);
      }
    }
    if (pick < 0) return false;   // no partner: the general step's fallback
    // one end occupied (or an electrode), the other empty (or an electrode):
    // flipping both bits and clearing the electrode sites' bits is the hop
    // whichever end is the source
    uint64_t nocc[C];
    bool back = true;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      nocc[c] = occ[c];
      if (C == 1 || (sel >> 6) == c) nocc[c] ^= 1ull << (sel & 63);
      if (C == 1 || (pick >> 6) == c) nocc[c] ^= 1ull << (pick & 63);
      nocc[c] &= accm[c];
      back = back && nocc[c] == B.key[c];
    }
    if (!back) return false;
#pragma unroll
    for (int c = 0; c < C; ++c) occ[c] = nocc[c];
    // electrode tallies on every lane (only the electrode lanes' are kept): the
    // rows are the sources when the walked set is the targets
    const int src = F.udst ? sel : pick, dst = F.udst ? pick : sel;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int s = lane + 64 * c;
      eo[c] += (s == dst ? 1 : 0) - (s == src ? 1 : 0);
    }
    if (lane == hl) dtv = lg_lane * F.rtot;
    ++h;
    if constexpr (PH) {
      ++ph[7];
      ++ph[8];
    }
    if (F.plast == -1) {
      // the first hop back from F since the entries changed: at a high miss
      // rate most pairs of states never come to a leap, so the bounds wait
      // for a second one
      F.plast = -2;
    } else if (F.plast < 0) {
      // the hop back from F is known: its row's and partner's bounds (maxima
      // over the sites before the row and before the partner)
      F.plast = pick;
      double lo = -INFINITY;
      float plo = -INFINITY;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (c > 0 && 64 * c >= S) break;
        const int s = lane + 64 * c;
        lo = fmax(lo, wave_max_f64(s < sel ? F.inc[c] : -INFINITY));
        plo = fmaxf(plo, wave_max_f32(s < pick && ((F.pnz[c] >> lane) & 1ull)
                                          ? F.pcum[c] : -INFINITY));
      }
      F.lo = in_vgpr(lo);
      F.plo = in_vgpr(plo);
      F.hi = in_vgpr(pick_lane_f64<C>(F.inc, sel));
      F.exs = in_vgpr(pick_lane_f64<C>(F.exc, sel));
      F.phi = in_vgpr(pick_lane_value<C>(F.pcum, pick));
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int s = lane + 64 * c;
        F.dlt[c] = (s == dst ? 1 : 0) - (s == src ? 1 : 0);
      }
    }
    mark(6);
    return true;
  };

  if constexpr (kFast) {
    // Hops alternate between the entries.  `one`: entry E1 holds (or, after a
    // miss, is to hold) the current state.  A special hop's work first, then fast
    // hops and leaps for as long as they apply (the current state's entry must
    // hold it), then a general step for the hop that stopped them.
    bool one = false;
    while (h < h_stop) {
      if (h == special && !special_work()) break;
      if (cur_ok) {
        if (one) {
          while (fast_hop(E1, E0))
            if (!fast_hop(E0, E1)) {
              one = false;
              break;
            }
        } else {
          while (fast_hop(E0, E1))
            if (!fast_hop(E1, E0)) {
              one = true;
              break;
            }
        }
        if (h >= h_lim) continue;   // the range's end or a special hop
      }
      if (!one) {
        if (!step(E0, E1)) break;
      } else {
        if (!step(E1, E0)) break;
      }
      one = !one;
    }
  } else {
    // general steps only, unrolled by the entries' roles
    if (h < h_stop)
      for (;;) {
        if (!step(E0, E1)) break;
        if (!step(E1, E0)) break;
      }
  }
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);   // a stopped trajectory leaves the loop raised
  flush_clock();
  if (!team.writer()) return;
  if constexpr (PH) {
    if (p.phases && lane == 0)
      for (int k = 0; k < FP32_PHASES; ++k)
        atomicAdd(reinterpret_cast<unsigned long long*>(p.phases + k), (unsigned long long)ph[k]);
  }
  if (lane == 0) {
    const int64_t dt_ticks = wall_clock64() - t_start;
    o->ticks = (int32_t)(dt_ticks < INT32_MAX ? dt_ticks : INT32_MAX);
    o->time = time;
    o->hops_done = h > prehops ? h - prehops : 0;
    o->status = status;
#pragma unroll
    for (int w = 0; w < OCCW; ++w) o->occ[w] = w < C ? occ[w] : 0ull;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int s = lane + 64 * c;
    if (s >= N && s < S) o->eo[s - N] = eo[c];
    if (REC && p.occ_time && s < N) p.occ_time[(size_t)t * NMAX + s] = ot[c];
  }
}

}  // namespace kmc
