"""CPU oracle for the kmc_dn hop loop -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, and only as the checker (or the timed CPU baseline).  The
product package ``kmc_dn_amd`` never imports it.

``liboracle.so`` (built from ``oracle/kmc_oracle.c`` by ``oracle/Makefile``)
holds the restatements:

* :func:`numba_simulate` -- ``_simulate_discrete_record``
  (kmc_dopant_networks.py:33-135) in FP64 with a replayed uniform stream.
* :func:`go_simulate` -- ``simulate`` / ``simulateRecordPlus``
  (goSimulation/simulation.go:194-325, :327-432) in FP32.
* :func:`philox4x32_10`, :func:`stream_uniforms` -- the engine's RNG contract.
* :func:`prob_simulate` -- ``probSimulate``
  (goSimulation/probabilitySimulation.go:53-157), the mean-field integrator.
* :func:`average_current` -- ``calcAverageCurrent`` (simulation.go:671-736).

``oracle.pyref`` repeats the numba restatement in pure Python for tiny cases
(a second, independent reading of the same reference lines).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# KMC_ORACLE_LIB: another build of the same source (the sanitizer build of
# tests/test_sanitizers.py)
LIB_PATH = os.environ.get("KMC_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.philox4x32_10.restype = None
        L.kmc_stream_fill.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, _dp]
        L.kmc_stream_fill.restype = None
        L.oracle_numba_simulate.argtypes = [
            C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
            _dp, _dp, _dp, _dp, _u8p, _i64p, C.c_int64, C.c_int, _dp, _dp,
            _dp, C.c_uint64, C.c_uint64, _i32p, _dp, _dp]
        L.oracle_numba_simulate.restype = C.c_int
        L.oracle_go_simulate.argtypes = [
            C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
            _dp, _dp, _dp, _dp, _u8p, _dp, C.c_int64, C.c_int, C.c_int, C.c_int, _dp, _dp,
            _dp, C.c_uint64, C.c_uint64, C.c_int64, _i32p, _dp, _i64p]
        L.oracle_go_simulate.restype = C.c_int
        L.oracle_numba_ensemble.argtypes = [
            C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
            _dp, _dp, _dp, _dp, _u8p, C.c_int64, _u64p, C.c_uint64, C.c_int64, C.c_int,
            _i64p, _dp, _u8p, _i32p]
        L.oracle_numba_ensemble.restype = C.c_int
        L.oracle_go_ensemble.argtypes = [
            C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
            _dp, _dp, _dp, _dp, _u8p, C.c_int64, _u64p, C.c_uint64, C.c_int64, C.c_int, C.c_int,
            C.c_int64, C.c_int, C.c_double, _dp, _dp, _u8p, _i32p, _i64p]
        L.oracle_go_ensemble.restype = C.c_int
        L.oracle_prob_simulate.argtypes = [
            C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
            _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64, C.c_int, _dp, _dp]
        L.oracle_prob_simulate.restype = C.c_double
        L.oracle_average_current.argtypes = [
            C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
            _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_average_current.restype = None
        L.oracle_max_threads.argtypes = []
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def philox4x32_10(counter, key):
    """One Philox4x32-10 block: 4 counter words, 2 key words -> 4 words."""
    c = np.ascontiguousarray(counter, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return o


def stream_uniforms(seed: int, hop0: int, n: int) -> np.ndarray:
    """Uniforms of the engine's stream: shape (n, 2) = (clock, event) per hop."""
    out = np.zeros(2 * n, dtype=np.float64)
    lib().kmc_stream_fill(seed, hop0, n, _p(out, _dp))
    return out.reshape(n, 2)


class NumbaResult:
    __slots__ = ("time", "occupation", "electrode_occupation", "site_energies", "traffic",
                 "occupations_in_time", "events", "totals", "status")


def numba_simulate(N, P, nu, kT, I_0, R, distances, E_constant, site_energies,
                   transitions_constant, occupation, electrode_occupation, hops, record=False,
                   uniforms=None, seed=0, hop0=0, log=False) -> NumbaResult:
    """Run the numba-loop restatement; inputs are copied, results returned."""
    S = N + P
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = np.array(site_energies, dtype=np.float64, copy=True)
    occ = np.array(occupation, dtype=np.uint8, copy=True)
    eo = np.array(electrode_occupation, dtype=np.int64, copy=True)
    traffic = np.zeros((S, S)) if record else None
    oit = np.zeros(N) if record else None
    u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64).ravel()
    if u is not None and u.size < 2 * hops:
        raise ValueError("need 2 uniforms per hop")
    ev = np.zeros(hops, dtype=np.int32) if log else None
    tot = np.zeros(hops, dtype=np.float64) if log else None
    t = C.c_double(0.0)
    st = lib().oracle_numba_simulate(
        N, P, nu, kT, I_0, R, _p(d, _dp), _p(E, _dp), _p(se, _dp), _p(tc, _dp), _p(occ, _u8p),
        _p(eo, _i64p), int(hops), int(bool(record)), _p(traffic, _dp), _p(oit, _dp), _p(u, _dp),
        int(seed), int(hop0), _p(ev, _i32p), _p(tot, _dp), C.byref(t))
    r = NumbaResult()
    r.time, r.occupation, r.electrode_occupation, r.site_energies = t.value, occ.astype(bool), eo, se
    r.traffic, r.occupations_in_time, r.events, r.totals, r.status = traffic, oit, ev, tot, st
    return r


class GoResult:
    __slots__ = ("time", "occupation", "electrode_occupation", "site_energies", "traffic",
                 "average_occupation", "events", "status", "cache_stats")


def go_simulate(N, P, nu, kT, I_0, R, distances, E_constant, site_energies, transitions_constant,
                occupation, hops, full_recompute=False, use_cache=False, record=False,
                prune_threshold=0.0, uniforms=None, seed=0, hop0=0, cache_max_bytes=1 << 28,
                log=False) -> GoResult:
    """Run the Go-loop restatement (FP32).  ``occupation`` is the start state."""
    S = N + P
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = np.array(site_energies, dtype=np.float64, copy=True)
    occ = np.array(occupation, dtype=np.uint8, copy=True)
    eo = np.zeros(P, dtype=np.float64)
    traffic = np.zeros((S, S)) if record else None
    avg = np.zeros(N) if record else None
    u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64).ravel()
    ev = np.zeros(2 * hops, dtype=np.int32) if log else None
    t = C.c_double(0.0)
    stats = np.zeros(3, dtype=np.int64)
    st = lib().oracle_go_simulate(
        N, P, float(prune_threshold), nu, kT, I_0, R, _p(d, _dp), _p(E, _dp), _p(tc, _dp),
        _p(se, _dp), _p(occ, _u8p), _p(eo, _dp), int(hops), int(bool(full_recompute)),
        int(bool(use_cache)), int(bool(record)), _p(traffic, _dp), _p(avg, _dp), _p(u, _dp),
        int(seed), int(hop0), int(cache_max_bytes), _p(ev, _i32p), C.byref(t), _p(stats, _i64p))
    r = GoResult()
    r.time, r.occupation, r.electrode_occupation, r.site_energies = t.value, occ.astype(bool), eo, se
    r.traffic, r.average_occupation, r.status = traffic, avg, st
    r.cache_stats = {"hits": int(stats[0]), "misses": int(stats[1]), "states": int(stats[2])}
    r.events = None if ev is None else ev.reshape(hops, 2)
    return r


def numba_ensemble(N, P, nu, kT, I_0, R, distances, E_constant, site_energies, transitions_constant,
                   occupation, seeds, hops, hop0=0, nthreads=0):
    """Independent trajectories (Philox key = seeds[k]) on OpenMP threads.

    Returns (electrode_occupation[K,P] int64, time[K], occupation[K,N], status[K]).
    """
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    K = seeds.size
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = np.ascontiguousarray(site_energies, dtype=np.float64)
    occ0 = np.ascontiguousarray(occupation, dtype=np.uint8)
    eo = np.zeros((K, P), dtype=np.int64)
    tm = np.zeros(K)
    occ = np.zeros((K, N), dtype=np.uint8)
    st = np.zeros(K, dtype=np.int32)
    lib().oracle_numba_ensemble(N, P, nu, kT, I_0, R, _p(d, _dp), _p(E, _dp), _p(se, _dp),
                                _p(tc, _dp), _p(occ0, _u8p), K, _p(seeds, _u64p), int(hop0),
                                int(hops), int(nthreads), _p(eo, _i64p), _p(tm, _dp),
                                _p(occ, _u8p), _p(st, _i32p))
    return eo, tm, occ, st


def go_ensemble(N, P, nu, kT, I_0, R, distances, E_constant, site_energies, transitions_constant,
                occupation, seeds, hops, full_recompute=False, use_cache=False,
                cache_max_bytes=1 << 26, hop0=0, nthreads=0, prune_threshold=0.0,
                cache_stats=None):
    """Go-restatement trajectories on OpenMP threads; same return layout.

    ``cache_stats``: optional int64 array (K, 3) that receives each trajectory's
    cache hits, misses and distinct occupation states."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    K = seeds.size
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = np.ascontiguousarray(site_energies, dtype=np.float64)
    occ0 = np.ascontiguousarray(occupation, dtype=np.uint8)
    eo = np.zeros((K, P), dtype=np.float64)
    tm = np.zeros(K)
    occ = np.zeros((K, N), dtype=np.uint8)
    st = np.zeros(K, dtype=np.int32)
    if cache_stats is not None:
        assert cache_stats.dtype == np.int64 and cache_stats.shape == (K, 3)
        assert cache_stats.flags.c_contiguous
    lib().oracle_go_ensemble(N, P, nu, kT, I_0, R, _p(d, _dp), _p(E, _dp), _p(se, _dp),
                             _p(tc, _dp), _p(occ0, _u8p), K, _p(seeds, _u64p), int(hop0),
                             int(hops), int(bool(full_recompute)), int(bool(use_cache)),
                             int(cache_max_bytes), int(nthreads), float(prune_threshold),
                             _p(eo, _dp), _p(tm, _dp), _p(occ, _u8p), _p(st, _i32p),
                             _p(cache_stats, _i64p))
    return eo, tm, occ, st


class ProbResult:
    pass


def prob_simulate(N, P, nu, kT, I_0, R, distances, E_constant, site_energies,
                  transitions_constant, occupation, hops, record=False) -> ProbResult:
    """Run the probSimulate restatement (FP64).  ``occupation`` is the fractional
    start state (wrapperSimulateProbability passes all 0.5)."""
    S = N + P
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = np.array(site_energies, dtype=np.float64, copy=True)
    occ = np.array(occupation, dtype=np.float64, copy=True)
    eo = np.zeros(P, dtype=np.float64)
    traffic = np.zeros((S, S)) if record else None
    avg = np.zeros(N) if record else None
    t = lib().oracle_prob_simulate(N, P, nu, kT, I_0, R, _p(occ, _dp), _p(d, _dp), _p(E, _dp),
                                   _p(tc, _dp), _p(eo, _dp), _p(se, _dp), int(hops),
                                   int(bool(record)), _p(traffic, _dp), _p(avg, _dp))
    r = ProbResult()
    r.time, r.occupation, r.electrode_occupation, r.site_energies = t, occ, eo, se
    r.traffic, r.average_occupation = traffic, avg
    return r


def average_current(N, P, nu, kT, I_0, R, distances, E_constant, site_energies,
                    transitions_constant, occupation) -> np.ndarray:
    """calcAverageCurrent restatement: currents (P,) at the mean occupation."""
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = np.ascontiguousarray(site_energies, dtype=np.float64)
    occ = np.ascontiguousarray(occupation, dtype=np.float64)
    out = np.zeros(P, dtype=np.float64)
    lib().oracle_average_current(N, P, nu, kT, I_0, R, _p(occ, _dp), _p(d, _dp), _p(E, _dp),
                                 _p(tc, _dp), _p(se, _dp), _p(out, _dp))
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())
