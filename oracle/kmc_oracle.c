/*
 * CPU oracle for the KMC hop loop of kmc_dn.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the CPU baseline.  The product
 * (kmc_dn_amd/) never links or calls it.
 *
 * Contents
 *   philox4x32_10      Philox4x32-10 counter-based RNG (Salmon et al., SC'11,
 *                      "Parallel random numbers: as easy as 1, 2, 3").  Pinned
 *                      by the Random123 known-answer vectors and by rocRAND's
 *                      philox4x32_10 engine (tests/test_oracle_philox.py).
 *   oracle_numba_simulate
 *                      Line-by-line restatement of _simulate_discrete_record,
 *                      kmc_dopant_networks.py:33-135 with _transition_possible
 *                      :137-163, FP64, dense S^2.  The numba RNG (MT19937) is
 *                      replaced by a uniform stream: u[2h] drives the clock,
 *                      u[2h+1] the event (the reference draws the exponential
 *                      first, :100, then the event uniform, :106).
 *   oracle_go_simulate Restatement of goSimulation/simulation.go:194-325
 *                      (simulate) and :327-432 (simulateRecordPlus): FP32
 *                      compacted pair list, O(N) makeJump update, problist
 *                      cache keyed by occupation (collision-free key here; the
 *                      reference's uint64 key is wrong for N > 64,
 *                      simulation.go:29-38).  Go's ziggurat ExpFloat64 and
 *                      Float32 are replaced by the same uniform stream.
 *   oracle_*_ensemble  OpenMP ensemble drivers (one trajectory per thread)
 *                      used as the CPU baseline.
 *   oracle_prob_simulate
 *                      Line-by-line restatement of probSimulate,
 *                      goSimulation/probabilitySimulation.go:53-157 with
 *                      calcProbTransitions :8-38 and probTransitionPossible
 *                      :40-50: the deterministic mean-field integrator behind
 *                      wrapperSimulateProbability (simulationWrapper.go:235-250),
 *                      FP64.
 *   oracle_average_current
 *                      Restatement of calcAverageCurrent, simulation.go:671-736,
 *                      with its FP32/FP64 mix (inputs cast to float32 as the Go
 *                      wrapper does, simulationWrapper.go:116-122).
 *
 * Build: oracle/Makefile, gcc -O3 -march=x86-64-v3 -ffp-contract=off.  Contraction
 * must stay off: numba (LLVM without fast-math) evaluates every a*b+c as a
 * rounded multiply followed by a rounded add.
 *
 * Parity status of each function is stated in its test; nothing here is
 * "parity unpinned" except the Go RNG streams (statistical pins only).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OK 0
#define ERR_ARGS 1
#define ERR_ZERO_TOTAL 2
#define ERR_NOMEM 3

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                        */
/* ------------------------------------------------------------------ */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* 53-bit uniform in [0,1) from two words (numpy's random_double layout). */
static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* The engine's stream contract: key = seed, counter = (hop lo, hop hi, 0, 0);
 * words 0,1 -> clock uniform, words 2,3 -> event uniform. */
void kmc_stream_uniforms(uint64_t seed, uint64_t hop, double* u_clock, double* u_event) {
    uint32_t ctr[4] = {(uint32_t)hop, (uint32_t)(hop >> 32), 0u, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    *u_clock = u53(o[0], o[1]);
    *u_event = u53(o[2], o[3]);
}

/* Fill 2*n uniforms (clock, event interleaved) for hops hop0 .. hop0+n-1. */
void kmc_stream_fill(uint64_t seed, uint64_t hop0, int64_t n, double* out) {
    for (int64_t h = 0; h < n; ++h)
        kmc_stream_uniforms(seed, hop0 + (uint64_t)h, &out[2 * h], &out[2 * h + 1]);
}

/* ------------------------------------------------------------------ */
/* numba path: kmc_dopant_networks.py:33-163                           */
/* ------------------------------------------------------------------ */

/* _transition_possible, kmc_dopant_networks.py:137-163 */
static inline int numba_possible(int64_t i, int64_t j, int64_t N, const uint8_t* occ) {
    if (i >= N && j >= N) return 0;
    else if (i >= N) return occ[j] ? 0 : 1;
    else if (j >= N) return occ[i] ? 1 : 0;
    else if (i == j) return 0;
    else return (!occ[i] || occ[j]) ? 0 : 1;
}

/*
 * _simulate_discrete_record, kmc_dopant_networks.py:33-135.
 * occupation (N bytes, 0/1), electrode_occupation (P) and site_energies[0:N]
 * are updated in place, as in the reference.  `uniforms` (2*hops) replays the
 * random stream; if NULL the Philox stream (seed, hop0 + h) is used.
 * events/totals (optional, length hops) log the event index k = from*S + to
 * and the total rate problist[-1] of every hop.
 * Returns 0, or ERR_ZERO_TOTAL where the reference raises ZeroDivisionError
 * (1/problist[-1] with problist[-1] == 0, :100).
 */
int oracle_numba_simulate(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                          const double* distances, const double* E_constant, double* site_energies,
                          const double* transitions_constant, uint8_t* occupation,
                          int64_t* electrode_occupation, int64_t hops, int record,
                          double* traffic, double* occupations_in_time,
                          const double* uniforms, uint64_t seed, uint64_t hop0,
                          int32_t* events, double* totals, double* time_out) {
    const int64_t S = N + P;
    if (N < 0 || P < 0 || S <= 0 || hops < 0) return ERR_ARGS;
    double* transitions = (double*)malloc(sizeof(double) * S * S);
    double* problist = (double*)malloc(sizeof(double) * S * S);
    if (!transitions || !problist) { free(transitions); free(problist); return ERR_NOMEM; }
    if (record) {
        if (traffic) memset(traffic, 0, sizeof(double) * S * S);                 /* :53 */
        if (occupations_in_time) memset(occupations_in_time, 0, sizeof(double) * N); /* :54 */
    }
    double time = 0.0;                                                          /* :55 */
    int status = OK;
    for (int64_t hop = 0; hop < hops; ++hop) {                                 /* :57 */
        /* site energies, :59-66 */
        for (int64_t i = 0; i < N; ++i) {
            double acceptor_interaction = 0.0;
            site_energies[i] = E_constant[i];
            for (int64_t j = 0; j < N; ++j)
                if (j != i)
                    acceptor_interaction += (double)(1 - (int)occupation[j]) / distances[i * S + j];
            site_energies[i] += -I_0 * R * acceptor_interaction;
        }
        /* transitions, :69-84 */
        for (int64_t i = 0; i < S; ++i)
            for (int64_t j = 0; j < S; ++j) {
                double t;
                if (!numba_possible(i, j, N, occupation)) {
                    t = 0.0;
                } else {
                    double dE;
                    if (i < N && j < N)
                        dE = site_energies[j] - site_energies[i] - I_0 * R / distances[i * S + j];
                    else
                        dE = site_energies[j] - site_energies[i];
                    if (dE > 0) t = nu * exp(-dE / kT);
                    else t = nu;
                }
                /* :87  transitions = transitions_constant*transitions */
                transitions[i * S + j] = transitions_constant[i * S + j] * t;
            }
        /* cumulative sum, :91-97 */
        for (int64_t k = 0; k < S * S; ++k)
            problist[k] = (k == 0) ? transitions[0] : transitions[k] + problist[k - 1];
        const double total = problist[S * S - 1];
        if (totals) totals[hop] = total;
        if (total == 0.0) { status = ERR_ZERO_TOTAL; break; }
        /* hop time, :100 -- numba: -log(1 - random()) * scale */
        double u_c, u_e;
        if (uniforms) { u_c = uniforms[2 * hop]; u_e = uniforms[2 * hop + 1]; }
        else kmc_stream_uniforms(seed, hop0 + (uint64_t)hop, &u_c, &u_e);
        const double scale = 1.0 / total;
        const double hop_time = -log(1.0 - u_c) * scale;
        /* normalise + first index with problist[i] >= event, :103-110 */
        int64_t event = -1;
        for (int64_t k = 0; k < S * S; ++k)
            if (problist[k] / total >= u_e) { event = k; break; }
        int64_t from, to;
        if (event < 0) { from = 0; to = 0; event = 0; }   /* no match: int(u/S), int(u%S) = 0, 0 */
        else { from = event / S; to = event % S; }
        if (events) events[hop] = (int32_t)event;
        /* perform hop, :116-123 */
        if (from < N) occupation[from] = 0;
        else electrode_occupation[from - N] -= 1;
        if (to < N) occupation[to] = 1;
        else electrode_occupation[to - N] += 1;
        /* record, :126-130 (waiting time attributed to the post-hop state) */
        if (record) {
            if (traffic) traffic[from * S + to] += 1;
            if (occupations_in_time)
                for (int64_t i = 0; i < N; ++i)
                    if (occupation[i]) occupations_in_time[i] += hop_time;
        }
        time += hop_time;                                                       /* :133 */
    }
    free(transitions);
    free(problist);
    if (time_out) *time_out = time;
    return status;
}

/* ------------------------------------------------------------------ */
/* Go path: goSimulation/simulation.go                                 */
/* ------------------------------------------------------------------ */

/* transition_possible, simulation.go:40-55 */
static inline int go_possible(int i, int j, int NSites, const uint8_t* occ) {
    if (i == j) return 0;
    if (i >= NSites && j >= NSites) return 0;
    else if (i >= NSites && occ[j]) return 0;
    else if (j >= NSites && !occ[i]) return 0;
    if (i < NSites && j < NSites && (!occ[i] || occ[j])) return 0;
    return 1;
}

/* getRandomEvent, simulation.go:163-188: bisect, then walk, to the first
 * probList[i] >= r.  `u` replaces rand.Float32(). */
static int go_random_event(const float* probList, int n, float u) {
    const float eventRand = u * probList[n - 1];
    int e_step = n / 2;
    int i = e_step;
    for (;;) {
        if (e_step >= 2) e_step = e_step / 2;
        if (probList[i] < eventRand) {
            i += e_step;
            if (i >= n) i = n - 1;
        } else if (i > 0 && probList[i - 1] >= eventRand) {
            i -= e_step;
            if (i < 0) i = 0;
        } else {
            return i;
        }
    }
}

/* Open-addressing map from occupation key (<= 4 words, N <= 256) to a stored
 * probList; replaces Go's map[uint64]*probabilities with a collision-free key. */
typedef struct {
    uint64_t key[4];
    int32_t count;   /* sightings (countProbs) */
    int32_t slot;    /* index into the problist pool, -1 if not stored */
} go_entry;

typedef struct {
    go_entry* tab;
    int64_t cap;      /* power of two */
    int64_t used;
    float* pool;
    int64_t pool_cap; /* number of storable problists */
    int64_t pool_used;
    int nt;           /* problist length */
} go_cache;

static uint64_t go_hash(const uint64_t k[4]) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int w = 0; w < 4; ++w) { h ^= k[w]; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 31; }
    return h;
}

static go_entry* go_cache_find(go_cache* c, const uint64_t k[4], int insert) {
    if (insert && (c->used + 1) * 2 > c->cap) {  /* grow */
        int64_t ncap = c->cap ? c->cap * 2 : 1024;
        go_entry* nt = (go_entry*)calloc((size_t)ncap, sizeof(go_entry));
        if (!nt) return NULL;
        for (int64_t i = 0; i < ncap; ++i) nt[i].count = -1;
        for (int64_t i = 0; i < c->cap; ++i) {
            if (c->tab[i].count < 0) continue;
            uint64_t p = go_hash(c->tab[i].key) & (uint64_t)(ncap - 1);
            while (nt[p].count >= 0) p = (p + 1) & (uint64_t)(ncap - 1);
            nt[p] = c->tab[i];
        }
        free(c->tab);
        c->tab = nt;
        c->cap = ncap;
    }
    if (!c->cap) return NULL;
    uint64_t p = go_hash(k) & (uint64_t)(c->cap - 1);
    for (;;) {
        go_entry* e = &c->tab[p];
        if (e->count < 0) {
            if (!insert) return NULL;
            memcpy(e->key, k, sizeof(e->key));
            e->count = 0;
            e->slot = -1;
            c->used++;
            return e;
        }
        if (!memcmp(e->key, k, sizeof(e->key))) return e;
        p = (p + 1) & (uint64_t)(c->cap - 1);
    }
}

/*
 * simulate (simulation.go:194-325) when full_recompute == 0, simulateRecordPlus
 * (:327-432) when full_recompute == 1.  FP64 inputs are narrowed to FP32 as
 * deFlattenFloatTo32/toFloat32 do (simulationWrapper.go:24-52).
 * use_cache: problist cache (record_problist).  For simulateRecordPlus the
 * number of stored lists is capped at 1.5e8/len(transitions) as in :353; for
 * simulate the reference has no cap, this restatement caps the pool at
 * cache_max_bytes.
 * occupation is used as given (the reference wrappers force it all-empty,
 * simulationWrapper.go:134-141; callers do that themselves).
 * electrode_occupation is FP64 counts, zeroed at start (:236-238).
 * traffic is net/antisymmetric (:310-311); average_occupation gets the
 * pre-hop state's waiting time (:312-316).
 */
int oracle_go_simulate(int64_t NSites, int64_t NElectrodes, double prune_threshold,
                       double nu_d, double kT_d, double I_0_d, double R_d,
                       const double* distances, const double* E_constant_d,
                       const double* transitions_constant, double* site_energies_d,
                       uint8_t* occupation, double* electrode_occupation, int64_t hops,
                       int full_recompute, int use_cache, int record,
                       double* traffic, double* average_occupation,
                       const double* uniforms, uint64_t seed, uint64_t hop0,
                       int64_t cache_max_bytes, int32_t* events_from_to, double* time_out,
                       int64_t* cache_stats) {
    const int N = (int)(NSites + NElectrodes);
    const int NS = (int)NSites;
    if (NSites < 0 || NElectrodes < 0 || N <= 0 || NS > 256) return ERR_ARGS;
    const float nu = (float)nu_d, kT = (float)kT_d, I_0 = (float)I_0_d, R = (float)R_d;
    float* d = (float*)malloc(sizeof(float) * N * N);
    float* tc = (float*)malloc(sizeof(float) * N * N);
    float* E = (float*)malloc(sizeof(float) * NS + 1);
    float* se = (float*)malloc(sizeof(float) * N);
    int* tf = (int*)malloc(sizeof(int) * N * N);
    int* tt = (int*)malloc(sizeof(int) * N * N);
    float* rate = (float*)malloc(sizeof(float) * N * N);
    float* prob = (float*)malloc(sizeof(float) * N * N);
    if (!d || !tc || !E || !se || !tf || !tt || !rate || !prob) {
        free(d); free(tc); free(E); free(se); free(tf); free(tt); free(rate); free(prob);
        return ERR_NOMEM;
    }
    for (int k = 0; k < N * N; ++k) { d[k] = (float)distances[k]; tc[k] = (float)transitions_constant[k]; }
    for (int i = 0; i < NS; ++i) E[i] = (float)E_constant_d[i];
    for (int i = 0; i < N; ++i) se[i] = (float)site_energies_d[i];

    /* transition list, :196-215 */
    float largest_tc = 0.0f;
    for (int k = 0; k < N * N; ++k) if (largest_tc < tc[k]) largest_tc = tc[k];
    const float cut = (float)prune_threshold * largest_tc;
    int nt = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            if (tc[i * N + j] > cut) { tf[nt] = i; tt[nt] = j; ++nt; }

    go_cache cache;
    memset(&cache, 0, sizeof(cache));
    cache.nt = nt;
    if (use_cache && nt > 0) {
        int64_t cap = full_recompute ? (int64_t)(150000000 / nt) : INT64_MAX;
        int64_t byte_cap = cache_max_bytes > 0 ? cache_max_bytes / ((int64_t)nt * 4) : 0;
        if (byte_cap < cap) cap = byte_cap;
        cache.pool_cap = cap;
        cache.pool = cap > 0 ? (float*)malloc((size_t)cap * nt * sizeof(float)) : NULL;
        if (cap > 0 && !cache.pool) cache.pool_cap = 0;
    }

    /* initial energies (simulate only), :226-234 */
    if (!full_recompute) {
        for (int i = 0; i < NS; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < NS; ++j)
                if (j != i && !occupation[j]) acc += 1.0f / d[i * N + j];
            se[i] = E[i] - I_0 * R * acc;
        }
    }
    for (int e = 0; e < (int)NElectrodes; ++e) electrode_occupation[e] = 0.0;
    double time = 0.0;
    uint32_t reuse_threshold = 1;
    uint64_t count_storage = 0, threshold_increase = 100000;
    int status = OK;
    int64_t n_hit = 0, n_miss = 0;

    for (int64_t hop = 0; hop < hops; ++hop) {
        const float* pl = NULL;
        go_entry* ent = NULL;
        uint64_t key[4] = {0, 0, 0, 0};
        if (use_cache) {
            for (int i = 0; i < NS; ++i) if (occupation[i]) key[i >> 6] |= 1ull << (i & 63);
            ent = go_cache_find(&cache, key, 0);
            if (ent && ent->slot >= 0) pl = cache.pool + (int64_t)ent->slot * nt;
            if (pl) ++n_hit; else ++n_miss;
        }
        if (!pl) {
            if (full_recompute) {   /* :378-386 */
                for (int i = 0; i < NS; ++i) {
                    float acc = 0.0f;
                    for (int j = 0; j < NS; ++j)
                        if (j != i && !occupation[j]) acc += 1.0f / d[i * N + j];
                    se[i] = E[i] - I_0 * R * acc;
                }
            }
            /* calcTransitionList, :58-80 */
            for (int k = 0; k < nt; ++k) {
                const int f = tf[k], t = tt[k];
                if (!go_possible(f, t, NS, occupation)) { rate[k] = 0.0f; continue; }
                float dE;
                if (f < NS && t < NS) dE = se[t] - se[f] - I_0 * R / d[f * N + t];
                else dE = se[t] - se[f];
                float r;
                if (dE > 0) r = nu * (float)exp((double)(-dE / kT));
                else r = nu;
                rate[k] = r * tc[f * N + t];
            }
            for (int k = 0; k < nt; ++k) prob[k] = k == 0 ? rate[0] : prob[k - 1] + rate[k];
            pl = prob;
            if (use_cache) {
                ent = go_cache_find(&cache, key, 1);
                if (ent) {
                    if (ent->count > 0) {
                        const uint32_t val = (uint32_t)ent->count;
                        ent->count++;
                        int store = full_recompute ? 1 : (val >= reuse_threshold);
                        if (store && ent->slot < 0 && cache.pool_used < cache.pool_cap) {
                            ent->slot = (int32_t)cache.pool_used++;
                            memcpy(cache.pool + (int64_t)ent->slot * nt, prob, sizeof(float) * nt);
                            if (!full_recompute && ++count_storage > threshold_increase) {
                                reuse_threshold++;
                                threshold_increase += 100000;
                            }
                        }
                    } else {
                        ent->count = 1;
                    }
                }
            }
        }
        const float total = nt > 0 ? pl[nt - 1] : 0.0f;
        if (!(total > 0.0f)) { status = ERR_ZERO_TOTAL; break; }
        double u_c, u_e;
        if (uniforms) { u_c = uniforms[2 * hop]; u_e = uniforms[2 * hop + 1]; }
        else kmc_stream_uniforms(seed, hop0 + (uint64_t)hop, &u_c, &u_e);
        /* :297 ExpFloat64()/total -> inversion of the stream's uniform */
        const double time_step = -log(1.0 - u_c) / (double)total;
        time += time_step;
        float uf = (float)u_e;
        if (uf >= 1.0f) uf = nextafterf(1.0f, 0.0f);   /* rand.Float32 never returns 1 */
        const int ev = go_random_event(pl, nt, uf);
        const int from = tf[ev], to = tt[ev];
        if (events_from_to) { events_from_to[2 * hop] = from; events_from_to[2 * hop + 1] = to; }
        if (record) {   /* :309-317 */
            if (traffic) { traffic[from * N + to] += 1; traffic[to * N + from] -= 1; }
            if (average_occupation)
                for (int i = 0; i < NS; ++i) if (occupation[i]) average_occupation[i] += time_step;
        }
        if (full_recompute) {   /* :418-429, no energy update */
            if (from < NS) occupation[from] = 0; else electrode_occupation[from - NS] -= 1.0;
            if (to < NS) occupation[to] = 1; else electrode_occupation[to - NS] += 1.0;
        } else {   /* makeJump, :107-130 */
            if (from < NS) {
                occupation[from] = 0;
                for (int j = 0; j < NS; ++j)
                    if (j != from) se[j] -= I_0 * R * (1.0f / d[j * N + from]);
            } else {
                electrode_occupation[from - NS] -= 1.0;
            }
            if (to < NS) {
                occupation[to] = 1;
                for (int j = 0; j < NS; ++j)
                    if (j != to) se[j] += I_0 * R * (1.0f / d[j * N + to]);
            } else {
                electrode_occupation[to - NS] += 1.0;
            }
        }
    }
    for (int i = 0; i < N; ++i) site_energies_d[i] = se[i];
    if (cache_stats) {   /* cache hits, misses, distinct occupation states seen */
        cache_stats[0] = n_hit;
        cache_stats[1] = n_miss;
        cache_stats[2] = cache.used;
    }
    free(cache.tab); free(cache.pool);
    free(d); free(tc); free(E); free(se); free(tf); free(tt); free(rate); free(prob);
    if (time_out) *time_out = time;
    return status;
}

/* ------------------------------------------------------------------ */
/* Ensemble drivers (CPU baseline): one trajectory per OpenMP thread.   */
/* ------------------------------------------------------------------ */

/* Each trajectory k starts from occupation0 (N bytes), uses Philox key
 * seeds[k] and counter hop0 + h, and writes eo_out[k*P..], time_out[k],
 * occ_out[k*N..] (final occupation) and status_out[k]. */
int oracle_numba_ensemble(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                          const double* distances, const double* E_constant,
                          const double* site_energies0, const double* transitions_constant,
                          const uint8_t* occupation0, int64_t ntraj, const uint64_t* seeds,
                          uint64_t hop0, int64_t hops, int nthreads,
                          int64_t* eo_out, double* time_out, uint8_t* occ_out, int32_t* status_out) {
    const int64_t S = N + P;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int64_t k = 0; k < ntraj; ++k) {
        double* se = (double*)malloc(sizeof(double) * S);
        uint8_t* occ = occ_out + k * N;
        memcpy(se, site_energies0, sizeof(double) * S);
        memcpy(occ, occupation0, (size_t)N);
        for (int64_t e = 0; e < P; ++e) eo_out[k * P + e] = 0;
        int st = oracle_numba_simulate(N, P, nu, kT, I_0, R, distances, E_constant, se,
                                       transitions_constant, occ, eo_out + k * P, hops, 0,
                                       NULL, NULL, NULL, seeds[k], hop0, NULL, NULL, &time_out[k]);
        status_out[k] = st;
        bad |= (st != 0);
        free(se);
    }
    return bad ? ERR_ZERO_TOTAL : OK;
}

/* prune_threshold as wrapperSimulatePruned (simulationWrapper.go:98-110);
 * cache_stats (optional): 3 per trajectory, as oracle_go_simulate writes them. */
int oracle_go_ensemble(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                       const double* distances, const double* E_constant,
                       const double* site_energies0, const double* transitions_constant,
                       const uint8_t* occupation0, int64_t ntraj, const uint64_t* seeds,
                       uint64_t hop0, int64_t hops, int full_recompute, int use_cache,
                       int64_t cache_max_bytes, int nthreads, double prune_threshold,
                       double* eo_out, double* time_out, uint8_t* occ_out, int32_t* status_out,
                       int64_t* cache_stats) {
    const int64_t S = N + P;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int64_t k = 0; k < ntraj; ++k) {
        double* se = (double*)malloc(sizeof(double) * S);
        uint8_t* occ = occ_out + k * N;
        memcpy(se, site_energies0, sizeof(double) * S);
        memcpy(occ, occupation0, (size_t)N);
        int st = oracle_go_simulate(N, P, prune_threshold, nu, kT, I_0, R, distances, E_constant,
                                    transitions_constant, se, occ, eo_out + k * P, hops,
                                    full_recompute, use_cache, 0, NULL, NULL, NULL, seeds[k], hop0,
                                    cache_max_bytes, NULL, &time_out[k],
                                    cache_stats ? cache_stats + 3 * k : NULL);
        status_out[k] = st;
        bad |= (st != 0);
        free(se);
    }
    return bad ? ERR_ZERO_TOTAL : OK;
}

/* ------------------------------------------------------------------ */
/* Mean-field estimators (probabilitySimulation.go, simulation.go)      */
/* ------------------------------------------------------------------ */
/* probTransitionPossible, probabilitySimulation.go:40-50 */
static inline double prob_possible(int64_t i, int64_t j, int64_t NSites, const double* occ) {
    if (i >= NSites && j >= NSites) return 0;
    if (i >= NSites) return 1 - occ[j];
    if (j >= NSites) return occ[i];
    return (1 - occ[j]) * occ[i];
}

/* probSimulate, probabilitySimulation.go:53-157.  Arrays are S x S row-major
 * (S = NSites + NElectrodes, the reference's N).  occupation and site_energies
 * are updated in place; electrode_occupation is zeroed, then accumulated;
 * traffic and average_occupation accumulate (record only).  Returns the time. */
double oracle_prob_simulate(int64_t NSites, int64_t NElectrodes, double nu, double kT, double I_0,
                            double R, double* occupation, const double* distances,
                            const double* E_constant, const double* transitions_constant,
                            double* electrode_occupation, double* site_energies, int64_t hops,
                            int record, double* traffic, double* average_occupation) {
    const int64_t N = NSites + NElectrodes;
    double* transitions = (double*)calloc((size_t)(N * N + 1), sizeof(double));
    double* difference = (double*)calloc((size_t)(NSites + 1), sizeof(double));
    if (!transitions || !difference) {
        free(transitions);
        free(difference);
        return NAN;
    }
    for (int64_t i = 0; i < NElectrodes; i++) electrode_occupation[i] = 0.0;
    double time = 0;
    for (int64_t hop = 0; hop < hops; hop++) {
        for (int64_t i = 0; i < NSites; i++) {
            difference[i] = 0;
            double acceptor_interaction = 0;
            for (int64_t j = 0; j < NSites; j++)
                if (j != i) acceptor_interaction += (1 - occupation[j]) / distances[i * N + j];
            site_energies[i] = E_constant[i] - I_0 * R * acceptor_interaction;
        }
        /* calcProbTransitions, :8-38 */
        double tot_rates = 0.0;
        for (int64_t i = 0; i < N; i++)
            for (int64_t j = 0; j < N; j++) {
                const double base_prob = prob_possible(i, j, NSites, occupation);
                double dE;
                if (i < NSites && j < NSites)
                    dE = site_energies[j] - site_energies[i] - I_0 * R / distances[i * N + j];
                else
                    dE = site_energies[j] - site_energies[i];
                double t;
                if (dE > 0)
                    t = base_prob * nu * exp(-dE / kT);
                else
                    t = base_prob * nu;
                t *= transitions_constant[i * N + j];
                transitions[i * N + j] = t;
                tot_rates += t;
                if (i < NSites) difference[i] -= t;
                if (j < NSites) difference[j] += t;
            }
        double max_change = 1.0;
        for (int64_t i = 0; i < NSites; i++) {
            const double newVal = occupation[i] + difference[i] / tot_rates;
            if (newVal < 0) {
                const double req_change = occupation[i] / -(difference[i] / tot_rates);
                if (req_change < max_change) max_change = req_change;
            }
            if (newVal > 1) {
                const double req_change = (1 - occupation[i]) / (difference[i] / tot_rates);
                if (req_change < max_change) max_change = req_change;
            }
        }
        const double time_step = 0.98 * (max_change) / tot_rates;
        time += time_step;
        if (record)
            for (int64_t i = 0; i < NSites; i++) average_occupation[i] += occupation[i] * time_step;
        for (int64_t i = 0; i < N; i++)
            for (int64_t j = 0; j < N; j++) {
                if (i >= NSites && j >= NSites) break;
                const double rate = transitions[i * N + j] * max_change / tot_rates;
                if (i < NSites)
                    occupation[i] -= rate;
                else
                    electrode_occupation[i - NSites] -= rate;
                if (j < NSites)
                    occupation[j] += rate;
                else
                    electrode_occupation[j - NSites] += rate;
                if (record) {
                    traffic[i * N + j] += rate;
                    traffic[j * N + i] -= rate;
                }
            }
        for (int64_t i = 0; i < NSites; i++) {
            if (occupation[i] < 0) occupation[i] = 0;
            if (occupation[i] > 1) occupation[i] = 1;
        }
    }
    free(transitions);
    free(difference);
    return time;
}

/* calcAverageCurrent, simulation.go:671-736, after the wrapper's float32 casts of
 * nu, kT, I_0, R, distances, E_constant, transitions_constant and site_energies
 * (only the electrode entries [NSites:] are read).  occupation is the
 * time-averaged occupation (float64); currents[e] = eoDifference / time_step. */
void oracle_average_current(int64_t NSites, int64_t NElectrodes, double nu64, double kT64,
                            double I064, double R64, const double* occupation,
                            const double* distances64, const double* E_constant64,
                            const double* tc64, const double* site_energies64,
                            double* currents) {
    const int64_t N = NSites + NElectrodes;
    const float nu = (float)nu64, kT = (float)kT64, I_0 = (float)I064, R = (float)R64;
    float* se = (float*)calloc((size_t)(N + 1), sizeof(float));
    double* transitions = (double*)calloc((size_t)(N * N + 1), sizeof(double));
    double* eoDifference = (double*)calloc((size_t)(NElectrodes + 1), sizeof(double));
    if (!se || !transitions || !eoDifference) {
        free(se);
        free(transitions);
        free(eoDifference);
        return;
    }
    for (int64_t i = 0; i < NSites; i++) {
        double acceptor_interaction = 0;
        for (int64_t j = 0; j < NSites; j++)
            if (j != i) acceptor_interaction += (1 - occupation[j]) / (double)(float)distances64[i * N + j];
        se[i] = (float)E_constant64[i] - I_0 * R * (float)acceptor_interaction;
    }
    for (int64_t e = 0; e < NElectrodes; e++) se[NSites + e] = (float)site_energies64[NSites + e];
    double tot_rates = 0.0;
    for (int64_t i = 0; i < N; i++)
        for (int64_t j = 0; j < N; j++) {
            const double base_prob = prob_possible(i, j, NSites, occupation);
            float dE;
            if (i < NSites && j < NSites)
                dE = se[j] - se[i] - I_0 * R / (float)distances64[i * N + j];
            else
                dE = se[j] - se[i];
            double t;
            if (dE > 0)
                t = base_prob * (double)nu * exp((double)(-dE / kT));
            else
                t = base_prob * (double)nu;
            t *= (double)(float)tc64[i * N + j];
            transitions[i * N + j] = t;
            tot_rates += t;
        }
    const double time_step = 0.98 / tot_rates;
    for (int64_t i = 0; i < N; i++)
        for (int64_t j = 0; j < N; j++) {
            if (i >= NSites && j >= NSites) break;
            const double rate = transitions[i * N + j] / tot_rates;
            if (i >= NSites) eoDifference[i - NSites] -= rate;
            if (j >= NSites) eoDifference[j - NSites] += rate;
        }
    for (int64_t e = 0; e < NElectrodes; e++) currents[e] = eoDifference[e] / time_step;
    free(se);
    free(transitions);
    free(eoDifference);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
