/*
 * kmc_dn_amd -- C-ABI of the MI355X KMC hop-loop engine (libkmc_dn_amd.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (KMC_OK == 0); on failure kmc_last_error() returns a thread-local message.
 * Nothing here aborts the process.
 *
 * Reference interfaces replaced (MUTUEL/kmc_dn):
 *   kmc_simulate_discrete   _simulate_discrete_record, kmc_dopant_networks.py:33-135
 *                           (called from makeSimulation, :582, :598-611)
 *   kmc_layout_* / kmc_eset_* / kmc_batch_*
 *                           parallelSimulations (goSimulation/simulationWrapper.go:274-315)
 *                           and channelSimulateRecord (:250-272): many independent
 *                           networks in one call, here as one device-resident batch
 *   wrapperSimulate, wrapperSimulateRecord, wrapperSimulateRecordPlus,
 *   wrapperSimulatePruned, parallelSimulations
 *                           exact-name exports with the cgo GoSlice signatures of
 *                           simulationWrapper.go:83-169, :274-315, so the reference's
 *                           ctypes callers (goSimulation/pythonBind.py:49-90,
 *                           goSimulation/parrallelSimulationBind.py:50-79) can load
 *                           this library unchanged.
 *   kmc_meanfield_simulate  probSimulate, goSimulation/probabilitySimulation.go:53-157
 *                           (behind wrapperSimulateProbability, simulationWrapper.go:235-250)
 *   kmc_average_current     calcAverageCurrent, goSimulation/simulation.go:671-736
 *                           (the estimator of wrapperSimulateCombined, :112-125)
 */
#ifndef KMC_DN_AMD_H
#define KMC_DN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KMC_NMAX 256 /* acceptors per network */
#define KMC_PMAX 16  /* electrodes per network */
#define KMC_OCCW 4   /* 64-bit words of an occupation bitmask */

enum {
  KMC_OK = 0,
  KMC_ERR_ARGS = 1,        /* bad sizes / pointers / ids */
  KMC_ERR_ZERO_RATE = 2,   /* total rate 0 or not finite (reference: ZeroDivisionError) */
  KMC_ERR_NOMEM = 3,       /* device or host allocation failed */
  KMC_ERR_DEVICE = 4,      /* HIP runtime error */
  KMC_ERR_UNSUPPORTED = 5, /* entry point / mode outside this engine's scope */
  KMC_ERR_LIMIT = 6,       /* network larger than the kernel's limits */
  KMC_ERR_TIMEOUT = 7,     /* kmc_batch_wait: the launch has not ended within the timeout */
  KMC_ERR_CANCELLED = 8    /* serving: a request cancelled or past its deadline (host loop) */
};

/* Precision / kernel of a batch run. */
enum {
  KMC_FP64_DENSE = 0,  /* reference arithmetic: dense S^2 FP64 rates, full energy
                          recompute, sequential row-major prefix sum */
  KMC_FP32_BATCHED = 1 /* FP32 log-domain rates, energies from a periodic FP64
                          recompute plus the Coulomb columns of the sites changed
                          since, FP64 wave scan + binary pick */
};

const char* kmc_last_error(void);
int kmc_version(void);

/* Selects the HIP device and creates the engine stream (idempotent). */
int kmc_init(int device);
/* Name, CU count, LDS bytes per workgroup of the current device. */
int kmc_device_info(char* name, int name_len, int* cu_count, int* lds_bytes);
/* Frees every layout, energy set and batch. */
int kmc_reset(void);

/*
 * _simulate_discrete_record (kmc_dopant_networks.py:33-135): `hops` hops of one
 * network with the reference's FP64 dense arithmetic, one trajectory per launch.
 * S = N + P.  distances, transitions_constant: S*S row-major; E_constant: N;
 * site_energies: S ([N:] = electrode voltages; [0:N] overwritten with the last
 * hop's energies); occupation: N bytes 0/1 (updated in place);
 * electrode_occupation: P (updated in place).  record != 0 fills traffic (S*S,
 * gross counts) and occupations_in_time (N, post-hop attribution), as the
 * reference does.  uniforms: 2*hops replayed uniforms (clock, event per hop) or
 * NULL for the Philox stream (seed, counter hop0 + h).  events_out / totals_out
 * (optional, length hops) log the event index from*S+to and problist[-1].
 */
int kmc_simulate_discrete(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                          const double* distances, const double* E_constant,
                          double* site_energies, const double* transitions_constant,
                          uint8_t* occupation, int64_t* electrode_occupation, int64_t hops,
                          int record, double* traffic, double* occupations_in_time,
                          const double* uniforms, uint64_t seed, uint64_t hop0,
                          int32_t* events_out, double* totals_out, double* time_out);

/*
 * No counterpart in the reference: kmc_simulate_discrete for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernel and results;
 * it calls no hipMalloc/hipFree and no allocating routine: every device buffer
 * is carved from one static scratch array per device (about 2.1 MiB, held
 * until the process ends).  Limits: S = N + P <= KMC_NOALLOC_MAX_SITES, and
 * with replayed uniforms or event/total logs at most KMC_NOALLOC_MAX_LOGGED_HOPS
 * hops (KMC_ERR_LIMIT beyond).  Calls on one device share its scratch and keep
 * the last call's contents; the engine lock serialises the calls of a process,
 * and calls that overlap on one device (bypassing it) would race.  The call is
 * synchronous: on return every output is written.
 */
#define KMC_NOALLOC_MAX_SITES 80
#define KMC_NOALLOC_MAX_LOGGED_HOPS 65536
int kmc_simulate_discrete_noalloc(int64_t N, int64_t P, double nu, double kT, double I_0,
                                  double R, const double* distances, const double* E_constant,
                                  double* site_energies, const double* transitions_constant,
                                  uint8_t* occupation, int64_t* electrode_occupation,
                                  int64_t hops, int record, double* traffic,
                                  double* occupations_in_time, const double* uniforms,
                                  uint64_t seed, uint64_t hop0, int32_t* events_out,
                                  double* totals_out, double* time_out);

/*
 * probSimulate (goSimulation/probabilitySimulation.go:53-157): the Go library's
 * deterministic mean-field integrator, `steps` steps of n networks at once
 * (network i: layout layouts[i], energy set esets[i] of that layout), FP64.
 * occupation: n*KMC_NMAX fractional occupations, updated in place (the reference
 * starts wrapperSimulateProbability from 0.5 everywhere); electrode_occupation:
 * n*KMC_PMAX fractional hop counts, zeroed then accumulated; site_energies
 * (n*KMC_NMAX, optional): acceptor energies of the last step; time (n,
 * optional); average_occupation (n*KMC_NMAX, optional): occupation x time;
 * traffic (optional): n blocks of Smax*Smax doubles (Smax = the largest S of the
 * n networks), each holding its network's S x S net traffic row-major from the
 * block start.  Sums run in another order than the reference's loops: results
 * agree to rounding.
 */
int kmc_meanfield_simulate(int32_t n, const int32_t* layouts, const int32_t* esets, int64_t steps,
                           double* occupation, double* electrode_occupation, double* site_energies,
                           double* time, double* average_occupation, double* traffic);

/*
 * calcAverageCurrent (goSimulation/simulation.go:671-736) of n networks at the
 * time-averaged occupations occupation (n*KMC_NMAX): currents (n*KMC_PMAX) =
 * each electrode's net inflow rate over the total rate, divided by the
 * reference's time step 0.98 / total.  Site energies and energy differences in
 * FP32, rates in FP64, as the reference computes them after its wrapper's
 * float32 casts.
 */
int kmc_average_current(int32_t n, const int32_t* layouts, const int32_t* esets,
                        const double* occupation, double* currents);

/* ---- device-resident batches (parallelSimulations, simulationWrapper.go:274-315) ---- */

/* Uploads one network geometry.  distances/transitions_constant: S*S row-major.
 * prune_threshold > 0 drops pairs with tc <= prune_threshold * max(tc) from the
 * FP32 kernel's pair set (wrapperSimulatePruned, simulation.go:199-215). */
int kmc_layout_create(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                      const double* distances, const double* transitions_constant,
                      double prune_threshold, int32_t* layout_out);

/* Attaches site coordinates (S*3, acceptors then electrodes) and the localisation
 * length ab to a layout, enabling the positions form (table_form 3).  The layout's
 * distances and transitions_constant must equal the reference closed forms
 * (kmc_dopant_networks.py:657-695, :824-830) within 1e-9 relative.  Layouts
 * created with prune_threshold > 0 are refused (KMC_ERR_UNSUPPORTED): the
 * positions form evaluates every pair from the coordinates. */
int kmc_layout_set_positions(int32_t layout, double ab, const double* positions);

/* Frees a layout's tables and every energy set created on it (ids may be reused). */
int kmc_layout_destroy(int32_t layout);

/* Uploads `count` energy sets of a layout: E_constant count*N, V count*P
 * (electrode voltages, i.e. site_energies[N:]).  Ids first..first+count-1. */
int kmc_eset_create(int32_t layout, int64_t count, const double* E_constant, const double* V,
                    int32_t* first_out);

/* Frees one energy set (its memory is released with the last set of its upload). */
int kmc_eset_destroy(int32_t eset);

/* Overwrites one energy set in place (E_constant N, V P doubles; no allocation).
 * Not ordered against running batches: rewrite only sets no queued run reads.
 * The reference recomputes E_constant on the host per voltage vector
 * (kmc_dopant_networks.py:865-899, update_V :798-821); this keeps it resident. */
int kmc_eset_write(int32_t eset, const double* E_constant, const double* V);

/* Overwrites n energy sets in one upload, ordered on the engine stream before
 * any later batch run: set ids[i] takes E_constant from E[i*KMC_NMAX ...] (its
 * layout's N values) and V from V[i*KMC_PMAX ...] (P values). */
int kmc_eset_write_many(int64_t n, const int32_t* ids, const double* E, const double* V);

typedef struct kmc_traj {
  int32_t layout;             /* from kmc_layout_create */
  int32_t eset;               /* from kmc_eset_create (same layout) */
  uint64_t seed;              /* Philox key */
  uint64_t hop0;              /* Philox counter of the first hop */
  int64_t prehops;            /* equilibration hops, not tallied */
  int64_t hops;               /* tallied hops */
  uint64_t occ[KMC_OCCW];     /* initial occupation bitmask, bit i = acceptor i */
} kmc_traj;

typedef struct kmc_run_opts {
  int32_t precision;          /* KMC_FP64_DENSE or KMC_FP32_BATCHED */
  int32_t waves_per_group;    /* FP32: trajectories per workgroup (1,2,4,8); 0 = auto */
  int32_t recompute_interval; /* FP32: hops between FP64 energy recomputes; 0 = 1024 */
  int32_t record;             /* bit 0: per-site occupation time; bit 1 (FP32): net
                                 traffic per ordered pair, Go convention */
  int32_t post_hop_record;    /* 1: numba convention (post-hop state), 0: pre-hop */
  int32_t grid_limit;         /* FP32: max resident workgroups; 0 = auto */
  int32_t use_graph;          /* 1: launch through a captured hipGraph */
  int32_t table_form;         /* FP32 pair terms: 0 auto, 1 {lt, lt+c} + c table in LDS
                                 (fewer VALU ops), 2 {lt, c} table (smaller LDS),
                                 3 computed from site positions (no table; needs
                                 kmc_layout_set_positions on every layout) */
  int32_t resume;             /* FP32: run each trajectory's [h_begin, h_end) set by
                                 kmc_batch_set_state, continuing from that state */
  int32_t team;               /* FP32: waves per trajectory, 1 (0) solo, 2 or 4: the
                                 walked set of each hop is split over the team (lower
                                 hop latency for batches that do not fill the GPU) */
  int32_t time_limit_us;      /* FP32 solo kernels: a trajectory that has run this many
                                 microseconds of device time stops at its next
                                 recompute point (a multiple of recompute_interval)
                                 with status KMC_OK and hops_done short of its range;
                                 a resume run from there continues bit-exactly.  Bounds
                                 the duration of a launch (schedulers, watchdogs).
                                 0 = none.  Refused with team kernels.  No reference
                                 counterpart. */
  int32_t state_store;        /* FP32 one-chunk solo kernels (S <= 64): LDS slots
                                 per trajectory for states evicted from its two-state
                                 cache, restored instead of re-evaluated when the
                                 trajectory returns to them.  0 = as many (at most 4)
                                 as fit without lowering the resident workgroups per
                                 CU, -1 = none, 1..4 = at most that many.  Changes no
                                 result, only the time per hop.  No reference
                                 counterpart (the Go library's occupation cache,
                                 simulation.go:254-296, keeps every state). */
} kmc_run_opts;

/* Uploads trajectory descriptors and allocates device outputs. */
int kmc_batch_create(const kmc_traj* trajs, int64_t ntraj, int32_t* batch_out);
/* Replaces the descriptors of a batch (ntraj up to its creation size; buffers
 * are reused, no allocation when it fits) -- for schedulers that refill slots. */
int kmc_batch_update(int32_t batch, const kmc_traj* trajs, int64_t ntraj);
/* Resume state per trajectory (caller order; any pointer may be NULL = keep):
 * absolute hop range [h_begin, h_end) of the next resume run, and the state it
 * continues from: time, eo[ntraj*KMC_PMAX], occ[ntraj*KMC_OCCW].  A range that
 * starts at a multiple of the recompute interval continues bit-exactly. */
int kmc_batch_set_state(int32_t batch, const int64_t* h_begin, const int64_t* h_end,
                        const double* time, const int64_t* eo, const uint64_t* occ);
/* Bucket launch plan for batches of uneven networks (FP32 batched path).
 * Each trajectory goes to the first bucket (in plan order) whose s_max >= its
 * layout's S and whose min_hops <= its budget (prehops + hops); every
 * trajectory must match one.  A bucket runs its trajectories in its own
 * work-queue kernel with its own pair-term form, workgroup size and team size;
 * the bucket kernels are launched as parallel branches of one stream fork/join
 * (one graph when use_graph is set).  A set plan overrides table_form,
 * waves_per_group and team of kmc_run_opts; nbuckets = 0 removes it.  No
 * reference counterpart: the reference runs one network per call
 * (simulationWrapper.go:83-169). */
#define KMC_MAX_BUCKETS 8
typedef struct kmc_bucket {
  int32_t s_max;              /* largest S of this bucket */
  int32_t table_form;         /* 1, 2 or 3 as in kmc_run_opts */
  int32_t waves_per_group;    /* solo kernels: 1, 2, 4 or 8 trajectories per workgroup */
  int32_t team;               /* waves per trajectory: 1 (solo), 2 or 4 */
  int64_t min_hops;           /* smallest budget this bucket takes */
  int32_t grid;               /* resident workgroups; 0 = as many as fit */
  int32_t s_min;              /* smallest S this bucket takes; 0 = no lower bound */
} kmc_bucket;
int kmc_batch_set_buckets(int32_t batch, const kmc_bucket* buckets, int32_t nbuckets);

/* Enqueues the batch on the engine stream (asynchronous). */
int kmc_batch_run(int32_t batch, const kmc_run_opts* opts);
/* Waits for the last run; elapsed_ms (optional) = HIP-event time of its kernels. */
int kmc_batch_sync(int32_t batch, double* elapsed_ms);
/* Host watchdog: waits at most timeout_ms for the last run to end, polling its
 * completion event without blocking the device.  KMC_OK (and elapsed_ms as in
 * kmc_batch_sync) when it ended, KMC_ERR_TIMEOUT when it had not; the launch is
 * left running either way (a later wait or sync still collects it; a launch
 * with kmc_run_opts.time_limit_us ends by itself).  No reference counterpart. */
int kmc_batch_wait(int32_t batch, double timeout_ms, double* elapsed_ms);
/* Copies results to host arrays (each may be NULL):
 * time[ntraj], eo[ntraj*KMC_PMAX], occ[ntraj*KMC_OCCW], status[ntraj],
 * hops_done[ntraj], occ_time[ntraj*KMC_NMAX] (only after a record run). */
int kmc_batch_results(int32_t batch, double* time, int64_t* eo, uint64_t* occ, int32_t* status,
                      int64_t* hops_done, double* occ_time);
/* Device time of each trajectory in the last FP32 run (caller order, seconds,
 * from the GPU's constant-rate counter): what a scheduler sizes hop ranges by. */
int kmc_batch_timing(int32_t batch, double* seconds);
/* Net traffic of the last run when it was an FP32 run with record bit 1 (fails
 * otherwise): traffic[ntraj*stride], trajectory j's S*S matrix at j*stride
 * (stride returned in *stride_out).  Batched FP64 runs refuse record bit 1. */
int kmc_batch_traffic(int32_t batch, double* traffic, int64_t* stride_out);
int kmc_batch_destroy(int32_t batch);

/* ---- exact-name exports of the reference's cgo library ----
 * Same symbols and by-value GoSlice signatures as goSimulation/simulationWrapper.go,
 * so goSimulation/pythonBind.py and parrallelSimulationBind.py can load this
 * library as libSimulation.so.  They return NaN (time) or -1 on error; see
 * kmc_last_error(). */
typedef struct kmc_GoSlice {
  double* data;
  int64_t len;
  int64_t cap;
} kmc_GoSlice;

/* simulationWrapper.go:83-96 (simulate, empty start, FP32) */
double wrapperSimulate(int64_t NSites, int64_t NElectrodes, double nu, double kT, double I_0,
                       double R, kmc_GoSlice occupation, kmc_GoSlice distances,
                       kmc_GoSlice E_constant, kmc_GoSlice transitions_constant,
                       kmc_GoSlice electrode_occupation, kmc_GoSlice site_energies,
                       int64_t hops, uint8_t record, kmc_GoSlice traffic,
                       kmc_GoSlice average_occupation);
/* simulationWrapper.go:98-110 (pair list pruned at prune_threshold * max tc) */
double wrapperSimulatePruned(int64_t NSites, int64_t NElectrodes, double prune_threshold,
                             double nu, double kT, double I_0, double R, kmc_GoSlice occupation,
                             kmc_GoSlice distances, kmc_GoSlice E_constant,
                             kmc_GoSlice transitions_constant, kmc_GoSlice electrode_occupation,
                             kmc_GoSlice site_energies, int64_t hops, uint8_t record,
                             kmc_GoSlice traffic, kmc_GoSlice average_occupation);
/* simulationWrapper.go:127-147 */
double wrapperSimulateRecord(int64_t NSites, int64_t NElectrodes, double nu, double kT,
                             double I_0, double R, kmc_GoSlice occupation, kmc_GoSlice distances,
                             kmc_GoSlice E_constant, kmc_GoSlice transitions_constant,
                             kmc_GoSlice electrode_occupation, kmc_GoSlice site_energies,
                             int64_t hops, uint8_t record, kmc_GoSlice traffic,
                             kmc_GoSlice average_occupation);
/* simulationWrapper.go:149-169 (records nothing) */
double wrapperSimulateRecordPlus(int64_t NSites, int64_t NElectrodes, double nu, double kT,
                                 double I_0, double R, kmc_GoSlice occupation,
                                 kmc_GoSlice distances, kmc_GoSlice E_constant,
                                 kmc_GoSlice transitions_constant,
                                 kmc_GoSlice electrode_occupation, kmc_GoSlice site_energies,
                                 int64_t hops, uint8_t record, kmc_GoSlice traffic,
                                 kmc_GoSlice average_occupation);
/* simulationWrapper.go:112-125: the FP32 hop loop from an empty network, then
 * calcAverageCurrent (simulation.go:671-736) at its time-averaged occupation;
 * electrode_occupation = those currents x time (same argument list as
 * wrapperSimulate). */
double wrapperSimulateCombined(int64_t, int64_t, double, double, double, double, kmc_GoSlice,
                               kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, kmc_GoSlice,
                               int64_t, uint8_t, kmc_GoSlice, kmc_GoSlice);
/* simulationWrapper.go:218-233: every acceptor at occupation 0.5, then
 * probSimulate (probabilitySimulation.go:53-157) for `hops` steps through
 * kmc_meanfield_simulate; occupation, site_energies[0:N] and
 * electrode_occupation are written back. */
double wrapperSimulateProbability(int64_t, int64_t, double, double, double, double, kmc_GoSlice,
                                  kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, kmc_GoSlice,
                                  int64_t, uint8_t, kmc_GoSlice, kmc_GoSlice);
/* simulationWrapper.go:171-216, a statistic of the Go occupation map (every
 * state a trajectory visits), which this engine does not keep (its kernels hold
 * the last two states only): returns NaN. */
double analyzeStateOverlap(int64_t, int64_t, double, double, double, double, kmc_GoSlice,
                           kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, kmc_GoSlice, int64_t,
                           uint8_t, kmc_GoSlice, kmc_GoSlice);
/* simulationWrapper.go:274-315: K networks in one call, all in one batch */
int64_t parallelSimulations(kmc_GoSlice NSites, kmc_GoSlice NElectrodes, kmc_GoSlice nu,
                            kmc_GoSlice kT, kmc_GoSlice I_0, kmc_GoSlice R,
                            kmc_GoSlice occupation, kmc_GoSlice distances,
                            kmc_GoSlice E_constant, kmc_GoSlice transitions_constant,
                            kmc_GoSlice electrode_occupation, kmc_GoSlice hops, kmc_GoSlice time,
                            kmc_GoSlice site_energies);

/* ---- measurement probes (no counterpart in the reference; used by bench.py) ---- */

/* Plain device-to-device copy of `bytes`, timed over `repeats` launches; the
 * kernel whose FETCH_SIZE/WRITE_SIZE calibrates rocprof's HBM byte counters. */
int kmc_probe_copy(int64_t bytes, int repeats, double* ms_per_copy);
/* LDS read bandwidth (bytes/s) of `blocks` x 256 threads streaming
 * width_bytes (4, 8 or 16) strided reads for `iters` x 8 rounds, or with
 * width_bytes 0 contiguous ds_read_b128 with immediate offsets, 16 per lane and
 * iteration (the LDS peak figure). */
int kmc_probe_lds(int width_bytes, int blocks, int iters, double* bytes_per_s);

/* Dependent-instruction latency of one wave in shader cycles per instruction:
 * kind 0 v_add_f64 (all lanes), 1 v_add_f64 (one lane), 2 v_add_f32, 3
 * v_fma_f64, 4 four interleaved independent v_add_f64 chains (issue rate). */
int kmc_probe_chain(int kind, int iters, double* cycles_per_instruction);

/* Runs the batch once on the FP64 dense kernel's phase-timing instantiation
 * (shader-clock reads at the phase boundaries of every hop, same work) and
 * returns trajectory 0's shader cycles summed per phase: cycles[0..6] =
 * energies, pair lists, rates, prefix chain, uniforms and clock, event search,
 * hop and records (n >= 7). */
int kmc_probe_fp64_phases(int32_t batch, uint64_t* cycles, int32_t n);

/* Runs the batch once as kmc_batch_run(batch, opts) would on the FP32 batched
 * path, with every launch on its kernel's phase-timing instantiation
 * (shader-clock reads at the phase boundaries of every hop), and returns the
 * shader cycles of each trajectory's first wave summed over trajectories:
 * cycles[0..6] = uniforms and FP64 recomputes, a miss's state-cache lookup,
 * energy update, electrode terms, row sums and team reduction, a miss's FP64
 * prefix scan and cache insert, clock and event selection, partner selection,
 * hop (occupation and tallies), a hit's lookup and copy; cycles[7] = the
 * number of state-cache hits, cycles[8] = the number of hops whose partner
 * prefix came from the state's entry (n >= 9).
 * Solo launches need a table form and 4 waves per workgroup, team launches
 * teams of 4 on 2 or 3 site chunks; no records, graph or resume.  The clock
 * reads wait for outstanding loads, so the instrumented hops run slower than
 * the ordinary kernel's. */
int kmc_probe_fp32_phases(int32_t batch, const kmc_run_opts* opts, uint64_t* cycles, int32_t n);

#ifdef __cplusplus
}
#endif
#endif /* KMC_DN_AMD_H */
