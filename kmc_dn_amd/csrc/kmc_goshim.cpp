// Exact-name exports of the reference's cgo library (goSimulation/simulationWrapper.go),
// so that its ctypes callers -- goSimulation/pythonBind.py:49-90 (callGoSimulation)
// and goSimulation/parrallelSimulationBind.py:50-79 (parrallelSimulation) -- can load
// libkmc_dn_amd.so in place of libSimulation.so without changes.
//
// Signatures follow the cgo export of each Go function: GoInt64/GoInt -> int64_t,
// float64 -> double, bool -> uint8_t, []float64 -> GoSlice {double* data; int64 len;
// int64 cap} passed by value.  Each call builds a device layout and energy set, runs
// one trajectory per network on the FP32 batched kernel (the Go library computes in
// FP32 too) and writes results back the way the Go code does:
//   * electrode_occupation: zeroed, then the float64 tallies (simulation.go:236-238);
//   * time: the return value (or time[] for parallelSimulations);
//   * wrapperSimulate/Record/Pruned with record: net traffic (:310-311) and pre-hop
//     occupation time (:312-316) accumulated into the caller's slices;
//   * the occupation slice is not written (the Go code never writes it back).
// Start states follow the Go wrappers: empty for wrapperSimulate* (:134-141), the
// given occupation (> 0 means occupied) for parallelSimulations (:256-263).
// wrapperSimulateProbability runs the mean-field integrator (kmc_meanfield_simulate)
// and wrapperSimulateCombined the hop loop plus calcAverageCurrent
// (kmc_average_current); analyzeStateOverlap, a diagnostic of the Go cache,
// returns NaN.
// Errors cannot be returned through these signatures: the functions return NaN
// (time) or -1 (parallelSimulations) and kmc_last_error() says why; the Go library
// would panic instead.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/kmc_dn_amd.h"

typedef kmc_GoSlice GoSlice;

namespace {

std::atomic<uint64_t> g_calls{0};

uint64_t next_seed() {
  // splitmix64 of a process-wide call counter: independent streams per call
  uint64_t z = (g_calls.fetch_add(1) + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

bool slice_ok(const GoSlice& s, int64_t need) { return s.data && s.len >= need; }

// One network, one trajectory; returns time or NaN.
double run_one(int64_t NSites, int64_t NElectrodes, double prune, double nu, double kT, double I_0,
               double R, const GoSlice& occupation, const GoSlice& distances,
               const GoSlice& E_constant, const GoSlice& transitions_constant,
               GoSlice electrode_occupation, const GoSlice& site_energies, int64_t hops_raw,
               bool record, GoSlice traffic, GoSlice average_occupation, bool from_empty,
               bool combined = false) {
  // The reference's ctypes binding passes hops as a 32-bit c_int to this 64-bit
  // argument; only the low 32 bits are defined, so read them as a signed int32.
  const int64_t hops = (int32_t)(uint32_t)hops_raw;
  const int64_t N = NSites, P = NElectrodes, S = N + P;
  if (kmc_init(-1) != KMC_OK) return NAN;
  if (N < 0 || P < 0 || hops < 0 || !slice_ok(distances, S * S) ||
      !slice_ok(transitions_constant, S * S) || !slice_ok(E_constant, N) ||
      !slice_ok(site_energies, S) || (P > 0 && !slice_ok(electrode_occupation, P)) ||
      (!from_empty && N > 0 && !slice_ok(occupation, N)))
    return NAN;
  int32_t lay = -1, es = -1, b = -1;
  if (kmc_layout_create(N, P, nu, kT, I_0, R, distances.data, transitions_constant.data, prune,
                        &lay) != KMC_OK)
    return NAN;
  double result = NAN;
  do {
    if (kmc_eset_create(lay, 1, E_constant.data, site_energies.data + N, &es) != KMC_OK) break;
    kmc_traj t;
    memset(&t, 0, sizeof t);
    t.layout = lay;
    t.eset = es;
    t.seed = next_seed();
    t.hops = hops;
    if (!from_empty)
      for (int64_t i = 0; i < N; ++i)
        if (occupation.data[i] > 0) t.occ[i >> 6] |= 1ull << (i & 63);
    if (kmc_batch_create(&t, 1, &b) != KMC_OK) break;
    kmc_run_opts o;
    memset(&o, 0, sizeof o);
    o.precision = KMC_FP32_BATCHED;
    o.record = record ? 3 : 0;
    if (kmc_batch_run(b, &o) != KMC_OK || kmc_batch_sync(b, nullptr) != KMC_OK) break;
    double time = 0;
    int64_t eo[KMC_PMAX];
    int32_t status = 0;
    std::vector<double> ot(record ? KMC_NMAX : 0);
    if (kmc_batch_results(b, &time, eo, nullptr, &status, nullptr, record ? ot.data() : nullptr) !=
        KMC_OK)
      break;
    if (status != KMC_OK) break;
    for (int64_t e = 0; e < P; ++e) electrode_occupation.data[e] = (double)eo[e];
    if (combined) {
      // simulateCombined (simulation.go:592-669): the hop loop's time-averaged
      // occupation through calcAverageCurrent (:671-736), times the elapsed time
      if (!(time > 0)) break;
      std::vector<double> avg(KMC_NMAX, 0.0), cur(KMC_PMAX, 0.0);
      for (int64_t i = 0; i < N; ++i) avg[i] = ot[i] / time;
      if (kmc_average_current(1, &lay, &es, avg.data(), cur.data()) != KMC_OK) break;
      for (int64_t e = 0; e < P; ++e) electrode_occupation.data[e] = cur[e] * time;
    }
    if (record) {
      int64_t stride = 0;
      if (kmc_batch_traffic(b, nullptr, &stride) != KMC_OK) break;
      std::vector<double> tr((size_t)stride);
      if (kmc_batch_traffic(b, tr.data(), &stride) != KMC_OK) break;
      if (slice_ok(traffic, S * S))
        for (int64_t k = 0; k < S * S; ++k) traffic.data[k] += tr[k];
      if (slice_ok(average_occupation, N))
        for (int64_t i = 0; i < N; ++i) average_occupation.data[i] += ot[i];
    }
    result = time;
  } while (false);
  if (b >= 0) kmc_batch_destroy(b);
  kmc_layout_destroy(lay);
  return result;
}

}  // namespace

extern "C" {

// simulationWrapper.go:83-96
double wrapperSimulate(int64_t NSites, int64_t NElectrodes, double nu, double kT, double I_0,
                       double R, GoSlice occupation, GoSlice distances, GoSlice E_constant,
                       GoSlice transitions_constant, GoSlice electrode_occupation,
                       GoSlice site_energies, int64_t hops, uint8_t record, GoSlice traffic,
                       GoSlice average_occupation) {
  return run_one(NSites, NElectrodes, 0.0, nu, kT, I_0, R, occupation, distances, E_constant,
                 transitions_constant, electrode_occupation, site_energies, hops, record != 0,
                 traffic, average_occupation, true);
}

// simulationWrapper.go:98-110
double wrapperSimulatePruned(int64_t NSites, int64_t NElectrodes, double prune_threshold, double nu,
                             double kT, double I_0, double R, GoSlice occupation, GoSlice distances,
                             GoSlice E_constant, GoSlice transitions_constant,
                             GoSlice electrode_occupation, GoSlice site_energies, int64_t hops,
                             uint8_t record, GoSlice traffic, GoSlice average_occupation) {
  return run_one(NSites, NElectrodes, prune_threshold, nu, kT, I_0, R, occupation, distances,
                 E_constant, transitions_constant, electrode_occupation, site_energies, hops,
                 record != 0, traffic, average_occupation, true);
}

// simulationWrapper.go:127-147 (the problist cache changes no result; it is not kept)
double wrapperSimulateRecord(int64_t NSites, int64_t NElectrodes, double nu, double kT, double I_0,
                             double R, GoSlice occupation, GoSlice distances, GoSlice E_constant,
                             GoSlice transitions_constant, GoSlice electrode_occupation,
                             GoSlice site_energies, int64_t hops, uint8_t record, GoSlice traffic,
                             GoSlice average_occupation) {
  return run_one(NSites, NElectrodes, 0.0, nu, kT, I_0, R, occupation, distances, E_constant,
                 transitions_constant, electrode_occupation, site_energies, hops, record != 0,
                 traffic, average_occupation, true);
}

// simulationWrapper.go:149-169 (simulateRecordPlus records nothing)
double wrapperSimulateRecordPlus(int64_t NSites, int64_t NElectrodes, double nu, double kT,
                                 double I_0, double R, GoSlice occupation, GoSlice distances,
                                 GoSlice E_constant, GoSlice transitions_constant,
                                 GoSlice electrode_occupation, GoSlice site_energies,
                                 int64_t hops, uint8_t record, GoSlice traffic,
                                 GoSlice average_occupation) {
  (void)record;
  return run_one(NSites, NElectrodes, 0.0, nu, kT, I_0, R, occupation, distances, E_constant,
                 transitions_constant, electrode_occupation, site_energies, hops, false, traffic,
                 average_occupation, true);
}

// simulationWrapper.go:112-125: the FP32 hop loop from an empty network, then
// calcAverageCurrent at its time-averaged occupation; electrode_occupation =
// those currents x time.  simulateCombined records traffic and occupation time
// whatever `record` says (simulation.go:646-654).
double wrapperSimulateCombined(int64_t NSites, int64_t NElectrodes, double nu, double kT,
                               double I_0, double R, GoSlice occupation, GoSlice distances,
                               GoSlice E_constant, GoSlice transitions_constant,
                               GoSlice electrode_occupation, GoSlice site_energies, int64_t hops,
                               uint8_t record, GoSlice traffic, GoSlice average_occupation) {
  (void)record;
  return run_one(NSites, NElectrodes, 0.0, nu, kT, I_0, R, occupation, distances, E_constant,
                 transitions_constant, electrode_occupation, site_energies, hops, true, traffic,
                 average_occupation, true, true);
}

// simulationWrapper.go:235-250: every acceptor at occupation 0.5 (written into the
// caller's slice, as the Go code does), then probSimulate
// (probabilitySimulation.go:53-157) for `hops` steps; occupation, site_energies
// [0:N] and electrode_occupation are written back, traffic and
// average_occupation accumulate when `record`.
double wrapperSimulateProbability(int64_t NSites, int64_t NElectrodes, double nu, double kT,
                                  double I_0, double R, GoSlice occupation, GoSlice distances,
                                  GoSlice E_constant, GoSlice transitions_constant,
                                  GoSlice electrode_occupation, GoSlice site_energies,
                                  int64_t hops_raw, uint8_t record, GoSlice traffic,
                                  GoSlice average_occupation) {
  const int64_t hops = (int32_t)(uint32_t)hops_raw;
  const int64_t N = NSites, P = NElectrodes, S = N + P;
  if (kmc_init(-1) != KMC_OK) return NAN;
  if (N < 0 || P < 0 || hops < 0 || !slice_ok(distances, S * S) ||
      !slice_ok(transitions_constant, S * S) || !slice_ok(E_constant, N) ||
      !slice_ok(site_energies, S) || (P > 0 && !slice_ok(electrode_occupation, P)) ||
      (N > 0 && !slice_ok(occupation, N)))
    return NAN;
  int32_t lay = -1, es = -1;
  if (kmc_layout_create(N, P, nu, kT, I_0, R, distances.data, transitions_constant.data, 0.0,
                        &lay) != KMC_OK)
    return NAN;
  double result = NAN;
  do {
    if (kmc_eset_create(lay, 1, E_constant.data, site_energies.data + N, &es) != KMC_OK) break;
    for (int64_t j = 0; j < N; ++j) occupation.data[j] = 0.5;
    std::vector<double> occ(KMC_NMAX, 0.0), eo(KMC_PMAX, 0.0), se(KMC_NMAX, 0.0);
    std::vector<double> avg(record ? KMC_NMAX : 0), tr(record ? (size_t)(S * S) : 0);
    for (int64_t j = 0; j < N; ++j) occ[j] = occupation.data[j];
    double time = 0.0;
    if (kmc_meanfield_simulate(1, &lay, &es, hops, occ.data(), eo.data(), se.data(), &time,
                               record ? avg.data() : nullptr, record ? tr.data() : nullptr) !=
        KMC_OK)
      break;
    for (int64_t j = 0; j < N; ++j) occupation.data[j] = occ[j];
    if (hops > 0)
      for (int64_t j = 0; j < N; ++j) site_energies.data[j] = se[j];
    for (int64_t e = 0; e < P; ++e) electrode_occupation.data[e] = eo[e];
    if (record) {
      if (slice_ok(traffic, S * S))
        for (int64_t q = 0; q < S * S; ++q) traffic.data[q] += tr[q];
      if (slice_ok(average_occupation, N))
        for (int64_t i = 0; i < N; ++i) average_occupation.data[i] += avg[i];
    }
    result = time;
  } while (false);
  kmc_layout_destroy(lay);
  return result;
}

// simulationWrapper.go:171-216 prints state-overlap statistics of the Go
// occupation map and returns 0: a diagnostic of a map of every visited state,
// which this engine does not keep (its kernels hold the last two states).
double analyzeStateOverlap(int64_t, int64_t, double, double, double, double, GoSlice, GoSlice,
                           GoSlice, GoSlice, GoSlice, GoSlice, int64_t, uint8_t, GoSlice, GoSlice) {
  return NAN;
}

// simulationWrapper.go:274-315: K networks, concatenated inputs, one call.  All
// networks run as one batch (one trajectory each) instead of one goroutine each.
int64_t parallelSimulations(GoSlice NSites, GoSlice NElectrodes, GoSlice nu, GoSlice kT,
                            GoSlice I_0, GoSlice R, GoSlice occupation, GoSlice distances,
                            GoSlice E_constant, GoSlice transitions_constant,
                            GoSlice electrode_occupation, GoSlice hops, GoSlice time,
                            GoSlice site_energies) {
  const int64_t K = NSites.len;
  if (kmc_init(-1) != KMC_OK) return -1;
  if (K <= 0 || NElectrodes.len < K || nu.len < K || kT.len < K || I_0.len < K || R.len < K ||
      hops.len < K || time.len < K || !NSites.data)
    return -1;
  std::vector<int32_t> lays;
  std::vector<kmc_traj> trajs(K);
  int64_t so = 0, eo_off = 0, comb = 0;
  bool ok = true;
  for (int64_t k = 0; k < K && ok; ++k) {
    const int64_t N = (int64_t)NSites.data[k], P = (int64_t)NElectrodes.data[k], S = N + P;
    if (N < 0 || P < 0 || occupation.len < so + N || E_constant.len < so + N ||
        distances.len < comb + S * S || transitions_constant.len < comb + S * S ||
        electrode_occupation.len < eo_off + P || site_energies.len < so + eo_off + S) {
      ok = false;
      break;
    }
    int32_t lay, es;
    if (kmc_layout_create(N, P, nu.data[k], kT.data[k], I_0.data[k], R.data[k],
                          distances.data + comb, transitions_constant.data + comb, 0.0,
                          &lay) != KMC_OK) {
      ok = false;
      break;
    }
    lays.push_back(lay);
    if (kmc_eset_create(lay, 1, E_constant.data + so, site_energies.data + so + eo_off + N, &es) !=
        KMC_OK) {
      ok = false;
      break;
    }
    kmc_traj& t = trajs[k];
    memset(&t, 0, sizeof t);
    t.layout = lay;
    t.eset = es;
    t.seed = next_seed();
    t.hops = (int64_t)hops.data[k];
    for (int64_t i = 0; i < N; ++i)
      if (occupation.data[so + i] > 0) t.occ[i >> 6] |= 1ull << (i & 63);
    so += N;
    eo_off += P;
    comb += S * S;
  }
  int32_t b = -1;
  if (ok) {
    kmc_run_opts o;
    memset(&o, 0, sizeof o);
    o.precision = KMC_FP32_BATCHED;
    ok = kmc_batch_create(trajs.data(), K, &b) == KMC_OK && kmc_batch_run(b, &o) == KMC_OK &&
         kmc_batch_sync(b, nullptr) == KMC_OK;
  }
  if (ok) {
    std::vector<double> t(K);
    std::vector<int64_t> eo((size_t)K * KMC_PMAX);
    std::vector<int32_t> st(K);
    ok = kmc_batch_results(b, t.data(), eo.data(), nullptr, st.data(), nullptr, nullptr) == KMC_OK;
    int64_t off = 0;
    for (int64_t k = 0; ok && k < K; ++k) {
      const int64_t P = (int64_t)NElectrodes.data[k];
      if (st[k] != KMC_OK) ok = false;
      time.data[k] = t[k];
      for (int64_t e = 0; e < P; ++e) electrode_occupation.data[off + e] = (double)eo[k * KMC_PMAX + e];
      off += P;
    }
  }
  if (b >= 0) kmc_batch_destroy(b);
  for (int32_t l : lays) kmc_layout_destroy(l);
  return ok ? 0 : -1;
}

}  // extern "C"
