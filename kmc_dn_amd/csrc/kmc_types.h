// Plain-old-data records shared by the host library and the kernels.
//
// A *layout* is one network geometry (acceptor + electrode positions and the
// constant tables derived from them); an *energy set* is one set of site
// constants E_constant[N] and electrode voltages V[P] on a layout; a
// *trajectory* is one independent KMC run (layout, energy set, RNG key, hop
// budget, initial occupation).
#pragma once
#include <stdint.h>

#include "../../include/kmc_dn_amd.h"

namespace kmc {

constexpr int NMAX = KMC_NMAX;   // acceptors per network
constexpr int PMAX = KMC_PMAX;   // electrodes per network
constexpr int OCCW = KMC_OCCW;   // 64-bit words of an occupation bitmask
constexpr int SMAX_FP32 = 192;   // sites of the FP32 batched kernel (3 lanes per site chunk)
constexpr int SMAX_FP64 = 80;    // sites of the FP64 dense kernel (LDS-resident lists)

struct LayoutDev {
  int32_t N, P, S, prune;
  double nu, kT, I0, R;
  double I0R;        // I_0 * R
  float scale2;      // log2(e) / kT: energies in kT units, base-2 exponent
  int32_t elw_ok;    // every electrode weight is finite (elw usable)
  // FP32 batched kernel.  tab32[u * S + s] = {lt, c} for the ordered pair
  // (s, u), symmetric in (s, u):
  //   lt = log2(nu * transitions_constant[s][u])  (-inf when impossible:
  //        diagonal, electrode-electrode, pruned or underflowed),
  //   c  = scale2 * I_0 * R / d[s][u] for acceptor-acceptor pairs, else 0.
  // 64 * 3 trailing entries {-inf, 0} let lanes past S read in bounds.
  const float2* tab32;
  // Same pairs as {lt, b = lt + c} (b = -inf where lt = -inf), and the c column
  // table ccol[u * S + s] = c(s, u) for acceptor columns u < N (N*S + 192 floats).
  const float2* tabB;
  const float* ccol;
  // Electrode weights elw[e * 192 + s] = nu * tc[s][N + e] for acceptor sites s
  // (0 for electrode sites, s >= N and pruned pairs; rows rounded up to a
  // multiple of 4, the padding zero): the
  // rate between acceptor s and electrode e is elw * min(1, 2^(e~_src - e~_dst)).
  const float* elw;
  const double* coul64;   // [N*N] I_0*R/d for acceptor pairs (symmetric, 0 diagonal)
  // Positions policy: site coordinates (w = 1 acceptor, 0 electrode) and the
  // closed forms lt = pos_l0 + pos_mkl * d, c = pos_kc / d (null if absent).
  const float4* pos4;
  float pos_l0, pos_mkl, pos_kc;
  int32_t has_pos;
  // FP64 dense kernel (reference arithmetic).
  const double* tc64;     // [S*S] transitions_constant
  const double* ird64;    // [S*S] (I_0*R)/d[i][j] (acceptor pairs; 0 elsewhere)
  const double* invd64;   // [N*N] 1.0/d[i][j] (0 on the diagonal)
  const double* dist64;   // [S*S] distances as given (mean-field estimators)
};



struct EsetDev {
  const double* E;   // [N] E_constant
  const double* V;   // [P] electrode voltages (site_energies[N:])
};

// Mean-field estimators of the Go library: probSimulate
// (goSimulation/probabilitySimulation.go:53-157) and calcAverageCurrent
// (goSimulation/simulation.go:671-736), one wave per network.
struct MeanfieldParams {
  const LayoutDev* layouts;
  const EsetDev* esets;
  const int32_t* layout_of;   // [n]
  const int32_t* eset_of;     // [n]
  double* occ;                // [n][NMAX] fractional occupation (in; probSimulate: out)
  double* eo;                 // [n][PMAX] electrode occupation (probSimulate) or currents
  double* se;                 // [n][NMAX] acceptor energies of the last step, or null
  double* time;               // [n] or null
  double* avg;                // [n][NMAX] occupation x time, or null
  double* traffic;            // [n][tstride] net traffic (S x S row-major), or null
  int64_t tstride;
  int64_t steps;
  int32_t n;
  int32_t pad;
};

struct TrajDev {
  int32_t layout;
  int32_t eset;
  uint64_t seed;      // Philox key
  uint64_t hop0;      // Philox counter of the first hop
  int64_t prehops;    // hops run before tallies start (not counted)
  int64_t hops;       // counted hops
  uint64_t occ[OCCW]; // initial occupation bitmask
  int64_t h_begin;    // absolute hop range of this launch (with BatchParams::resume)
  int64_t h_end;
};

struct TrajOut {
  double time;
  int64_t hops_done;
  int32_t status;
  int32_t ticks;     // steady-counter ticks of the trajectory's last run (saturated)
  uint64_t occ[OCCW];
  int64_t eo[PMAX];
};

struct GroupDev {
  int32_t layout;
  int32_t first;   // first trajectory (trajectories of a group are contiguous)
  int32_t count;   // <= waves per workgroup
  int32_t pad;
};

struct BatchParams {
  const TrajDev* trajs;
  const LayoutDev* layouts;
  const EsetDev* esets;
  const GroupDev* groups;
  int32_t ngroups;
  int32_t recompute_interval;   // hops between FP64 energy recomputes (>0)
  int32_t* queue;               // work-queue head, zero before launch
  TrajOut* out;
  double* occ_time;             // [ntraj * NMAX] or null
  double* traffic;              // [ntraj * traffic_stride] net traffic, or null
  int64_t traffic_stride;       // >= S*S of every layout in the batch
  int32_t post_hop_record;      // 1: numba convention (post-hop state)
  int32_t resume;               // 1: run [h_begin, h_end) from the state in `out`
  uint64_t* phases;             // [FP32_PHASES] sums over trajectories (the phase-timing
                                // instantiations only), or null
  int64_t time_limit;           // constant-rate clock ticks a solo trajectory runs before
                                // it stops at its next recompute point; 0: none
  int32_t store_slots;          // slots of each wave's LDS state store (one-chunk solo
                                // trajectories; 0: none)
  int32_t pad2;
};

// The state store of a one-chunk solo trajectory: states evicted from its two
// register entries, kept in LDS per wave.  Layout: STORE_MAX_SLOTS occupation
// keys (u64), then STORE_MAX_SLOTS nonzero-row lane masks (u64), then per slot
// the FP64 inclusive and exclusive row prefixes and the energies (FP32 values
// held as FP64, so that every array of a slot has the same per-lane stride) of
// the 64 lanes.
constexpr int STORE_MAX_SLOTS = 4;
constexpr int STORE_HEADER_BYTES = 2 * STORE_MAX_SLOTS * 8;
constexpr int STORE_SLOT_BYTES = 64 * (8 + 8 + 8);
constexpr int store_wave_bytes(int slots) {
  return slots > 0 ? STORE_HEADER_BYTES + slots * STORE_SLOT_BYTES : 0;
}

// Hop phases of the FP32 batched kernels timed by their phase-timing
// instantiations (shader cycles of each trajectory's first wave, summed over
// trajectories): 0 uniforms and FP64 recomputes (and fast-hop attempts that
// fell through to the general step), 1 energy update, electrode terms, row
// sums and team reduction of a miss, 2 FP64 prefix scan and cache insert of a
// miss, 3 clock and event selection, 4 partner selection, 5 hop (occupation,
// tallies) of the general step, 6 fast hops and leaps; slot 7 counts the hops
// from a cached state, slot 8 those whose partner prefix came from the state's
// entry, slot 9 those taken by leaps, slot 10 the new states restored from the
// LDS state store instead of evaluated.
constexpr int FP32_PHASES = 11;

struct PlainParams {
  const TrajDev* trajs;
  const LayoutDev* layouts;
  const EsetDev* esets;
  TrajOut* out;
  const double* uniforms;   // replayed stream (2 per hop) for trajectory 0, or null
  int32_t* events;          // [prehops+hops] event index from*S+to, or null
  double* totals;           // [prehops+hops] problist[-1] per hop, or null
  double* traffic;          // [ntraj * S*S] gross traffic, or null
  double* occ_time;         // [ntraj * NMAX] post-hop occupation time, or null
  double* se_out;           // [ntraj * NMAX] final site_energies[0:N], or null
  uint64_t* phases;         // [PLAIN_PHASES] shader cycles per hop phase of trajectory 0
                            // (the phase-timing instantiation only), or null
};

// Hop phases of the FP64 dense kernel timed by its phase-timing instantiation:
// energies, pair lists, rates, prefix chain, uniforms and clock, event search,
// hop and records.
constexpr int PLAIN_PHASES = 7;

// Status codes written per trajectory (also returned by the C-ABI).
enum : int32_t {
  ST_OK = KMC_OK,
  ST_ZERO_RATE = KMC_ERR_ZERO_RATE,
  ST_BAD_INPUT = KMC_ERR_ARGS,
};

}  // namespace kmc
