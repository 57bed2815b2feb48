"""Python view of the device engine: layouts, energy sets, batches.

This mirrors what the reference does with ``parrallelSimulation``
(goSimulation/parrallelSimulationBind.py:9-79) -- many independent networks in
one native call -- but keeps everything resident on the GPU: a :class:`Layout`
(geometry tables) is uploaded once, energy sets (E_constant, electrode
voltages) are uploaded per operating point, and a :class:`Batch` of
trajectories is launched, timed and read back as numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

TRAJ_DTYPE = np.dtype([("layout", np.int32), ("eset", np.int32), ("seed", np.uint64),
                       ("hop0", np.uint64), ("prehops", np.int64), ("hops", np.int64),
                       ("occ", np.uint64, (L.OCCW,))])
assert TRAJ_DTYPE.itemsize == C.sizeof(L.kmc_traj)


def init(device: int = -1) -> None:
    L.check(L.load().kmc_init(device))


def device_info() -> dict:
    lib = L.load()
    name = C.create_string_buffer(256)
    cu, lds = C.c_int(), C.c_int()
    L.check(lib.kmc_device_info(name, 256, C.byref(cu), C.byref(lds)))
    return {"name": name.value.decode(), "cu_count": cu.value, "lds_bytes": lds.value}


def reset() -> None:
    L.check(L.load().kmc_reset())


def occupation_to_bits(occ) -> np.ndarray:
    """Boolean occupation (..., N) -> (..., OCCW) uint64 bitmask words."""
    occ = np.asarray(occ, dtype=bool)
    N = occ.shape[-1]
    if N > L.NMAX:
        raise ValueError(f"N={N} exceeds {L.NMAX}")
    pad = np.zeros(occ.shape[:-1] + (L.OCCW * 64,), dtype=bool)
    pad[..., :N] = occ
    b = pad.reshape(occ.shape[:-1] + (L.OCCW, 64)).astype(np.uint64)
    return (b << np.arange(64, dtype=np.uint64)).sum(axis=-1, dtype=np.uint64)


def bits_to_occupation(bits, N: int) -> np.ndarray:
    bits = np.asarray(bits, dtype=np.uint64)
    x = (bits[..., :, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)
    return x.reshape(bits.shape[:-1] + (L.OCCW * 64,))[..., :N].astype(bool)


class Layout:
    """One network geometry on the device (kmc_layout_create)."""

    def __init__(self, N: int, P: int, nu: float, kT: float, I_0: float, R: float,
                 distances, transitions_constant, prune_threshold: float = 0.0):
        self.N, self.P, self.S = int(N), int(P), int(N) + int(P)
        self.nu, self.kT, self.I_0, self.R = float(nu), float(kT), float(I_0), float(R)
        d = np.ascontiguousarray(distances, dtype=np.float64)
        tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
        if d.shape != (self.S, self.S) or tc.shape != (self.S, self.S):
            raise ValueError(f"tables must be {self.S}x{self.S}")
        lid = C.c_int32()
        L.check(L.load().kmc_layout_create(self.N, self.P, self.nu, self.kT, self.I_0, self.R,
                                           L.ptr(d, L._dp), L.ptr(tc, L._dp),
                                           float(prune_threshold), C.byref(lid)))
        self.id = lid.value

    def energy_sets(self, E_constant, V) -> np.ndarray:
        """Upload energy sets (rows of E_constant [k,N] and V [k,P]); returns their ids."""
        E = np.ascontiguousarray(np.atleast_2d(E_constant), dtype=np.float64)
        Vv = np.ascontiguousarray(np.asarray(V, dtype=np.float64).reshape(E.shape[0], self.P))
        if E.shape[1] != self.N:
            raise ValueError("E_constant rows must have N entries")
        first = C.c_int32()
        L.check(L.load().kmc_eset_create(self.id, E.shape[0], L.ptr(E, L._dp), L.ptr(Vv, L._dp),
                                         C.byref(first)))
        return np.arange(first.value, first.value + E.shape[0], dtype=np.int32)

    def set_positions(self, ab: float, positions) -> "Layout":
        """Site coordinates (S, 3) for the positions form (table_form=3)."""
        xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.S, 3)
        L.check(L.load().kmc_layout_set_positions(self.id, float(ab), L.ptr(xyz, L._dp)))
        return self

    def write_energy_set(self, eset: int, E_constant, V) -> None:
        """Overwrite energy set `eset` of this layout in place (no allocation)."""
        E = np.ascontiguousarray(E_constant, dtype=np.float64).reshape(self.N)
        Vv = np.ascontiguousarray(V, dtype=np.float64).reshape(self.P)
        L.check(L.load().kmc_eset_write(int(eset), L.ptr(E, L._dp), L.ptr(Vv, L._dp)))

    @staticmethod
    def write_energy_sets(ids, E_pad, V_pad) -> None:
        """Overwrite energy sets ids[i] from rows of E_pad (n, NMAX) and V_pad (n, PMAX)
        (each set reads its layout's N and P leading entries) in one upload."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        E = np.ascontiguousarray(E_pad, dtype=np.float64).reshape(ids.size, L.NMAX)
        V = np.ascontiguousarray(V_pad, dtype=np.float64).reshape(ids.size, L.PMAX)
        L.check(L.load().kmc_eset_write_many(ids.size, L.ptr(ids, L._i32p), L.ptr(E, L._dp),
                                             L.ptr(V, L._dp)))

    @staticmethod
    def drop_energy_sets(ids) -> None:
        for i in np.atleast_1d(ids):
            L.check(L.load().kmc_eset_destroy(int(i)))

    def close(self) -> None:
        """Free the device tables and every energy set of this layout."""
        if self.id >= 0:
            L.check(L.load().kmc_layout_destroy(self.id))
            self.id = -1


@dataclass
class BatchResult:
    time: np.ndarray          # [K]
    eo: np.ndarray            # [K, PMAX] int64 electrode tallies
    occ_bits: np.ndarray      # [K, OCCW]
    status: np.ndarray        # [K]
    hops_done: np.ndarray     # [K]
    occ_time: np.ndarray | None = None   # [K, NMAX]

    def currents(self, P: int) -> np.ndarray:
        return self.eo[:, :P] / self.time[:, None]


def make_trajs(n: int, layout: int | np.ndarray, eset: int | np.ndarray, seeds, hops,
               occupation, prehops=0, hop0=0) -> np.ndarray:
    t = np.zeros(n, dtype=TRAJ_DTYPE)
    t["layout"] = layout
    t["eset"] = eset
    t["seed"] = seeds
    t["hop0"] = hop0
    t["prehops"] = prehops
    t["hops"] = hops
    t["occ"] = occupation_to_bits(occupation) if np.asarray(occupation).dtype == bool \
        else occupation
    return t


class Batch:
    """A device-resident set of trajectories (kmc_batch_create/run/sync/results)."""

    def __init__(self, trajs: np.ndarray):
        trajs = np.ascontiguousarray(trajs, dtype=TRAJ_DTYPE)
        self.n = trajs.shape[0]
        bid = C.c_int32()
        L.check(L.load().kmc_batch_create(trajs.ctypes.data_as(C.POINTER(L.kmc_traj)), self.n,
                                          C.byref(bid)))
        self.id = bid.value
        self.recorded = False

    def run(self, precision: int = L.FP32_BATCHED, waves_per_group: int = 0,
            recompute_interval: int = 0, record: bool = False, post_hop_record: bool = False,
            grid_limit: int = 0, use_graph: bool = False, traffic: bool = False,
            table_form: int = 0, resume: bool = False, team: int = 1,
            time_limit_us: int = 0, state_store: int = 0) -> None:
        rec = (1 if record else 0) | (2 if traffic else 0)
        o = L.kmc_run_opts(int(precision), int(waves_per_group), int(recompute_interval),
                           rec, int(post_hop_record), int(grid_limit), int(use_graph),
                           int(table_form), int(bool(resume)), int(team), int(time_limit_us),
                           int(state_store))
        L.check(L.load().kmc_batch_run(self.id, C.byref(o)))
        self.recorded = bool(record)

    def set_buckets(self, buckets) -> None:
        """Bucket plan: entries (s_max, table_form, waves_per_group[, team[, min_hops[,
        grid[, s_min]]]]); a trajectory goes to the first entry with s_min <= S <=
        s_max and budget >= min_hops.  Each bucket runs in its own kernel, all in one
        fork/join.  [] clears."""
        arr = (L.kmc_bucket * max(1, len(buckets)))()
        for i, b in enumerate(buckets):
            b = [int(x) for x in b] + [0] * (7 - len(b))
            s_max, form, w, team, min_hops, grid, s_min = b[:7]
            arr[i] = L.kmc_bucket(s_max, form, w, max(1, team), min_hops, grid, s_min)
        L.check(L.load().kmc_batch_set_buckets(self.id, arr, len(buckets)))

    def update(self, trajs: np.ndarray) -> None:
        """Replace the descriptors (reusing device buffers when they fit)."""
        trajs = np.ascontiguousarray(trajs, dtype=TRAJ_DTYPE)
        L.check(L.load().kmc_batch_update(self.id, trajs.ctypes.data_as(C.POINTER(L.kmc_traj)),
                                          trajs.shape[0]))
        self.n = trajs.shape[0]

    def set_state(self, h_begin=None, h_end=None, time=None, eo=None, occ_bits=None) -> None:
        """Hop ranges and start state of the next resume run (caller order)."""
        def a(x, dt, shape=None):
            if x is None:
                return None
            x = np.ascontiguousarray(x, dtype=dt)
            return x.reshape(shape) if shape else x
        hb = a(h_begin, np.int64)
        he = a(h_end, np.int64)
        tm = a(time, np.float64)
        e = a(eo, np.int64, (self.n, L.PMAX))
        o = a(occ_bits, np.uint64, (self.n, L.OCCW))
        L.check(L.load().kmc_batch_set_state(self.id, L.ptr(hb, L._i64p), L.ptr(he, L._i64p),
                                             L.ptr(tm, L._dp), L.ptr(e, L._i64p),
                                             L.ptr(o, L._u64p)))

    def sync(self) -> float:
        ms = C.c_double()
        L.check(L.load().kmc_batch_sync(self.id, C.byref(ms)))
        return ms.value

    def wait(self, timeout_ms: float) -> float | None:
        """Watchdog wait: the kernel time (ms) when the last run ended within
        timeout_ms, None when it had not (it keeps running)."""
        ms = C.c_double()
        rc = L.load().kmc_batch_wait(self.id, float(timeout_ms), C.byref(ms))
        if rc == L.ERR_TIMEOUT:
            return None
        L.check(rc)
        return ms.value

    def results(self) -> BatchResult:
        n = self.n
        t = np.zeros(n)
        eo = np.zeros((n, L.PMAX), dtype=np.int64)
        occ = np.zeros((n, L.OCCW), dtype=np.uint64)
        st = np.zeros(n, dtype=np.int32)
        hd = np.zeros(n, dtype=np.int64)
        ot = np.zeros((n, L.NMAX)) if self.recorded else None
        L.check(L.load().kmc_batch_results(self.id, L.ptr(t, L._dp), L.ptr(eo, L._i64p),
                                           L.ptr(occ, L._u64p), L.ptr(st, L._i32p),
                                           L.ptr(hd, L._i64p), L.ptr(ot, L._dp)))
        return BatchResult(t, eo, occ, st, hd, ot)

    def timing(self) -> np.ndarray:
        """Device seconds each trajectory took in the last FP32 run."""
        sec = np.zeros(self.n)
        L.check(L.load().kmc_batch_timing(self.id, L.ptr(sec, L._dp)))
        return sec

    def traffic_flat(self) -> np.ndarray:
        """Net traffic of the last run (an FP32 run with traffic=True), (n, stride):
        row j holds trajectory j's S_j x S_j matrix row-major from its start."""
        stride = C.c_int64()
        L.check(L.load().kmc_batch_traffic(self.id, None, C.byref(stride)))
        flat = np.zeros((self.n, stride.value))
        L.check(L.load().kmc_batch_traffic(self.id, L.ptr(flat, L._dp), C.byref(stride)))
        return flat

    def traffic(self, sizes) -> list:
        """Net traffic of the last FP32 run with traffic=True: one S_j x S_j matrix per
        trajectory j, ``sizes[j]`` = S_j of its layout."""
        flat = self.traffic_flat()
        sizes = np.broadcast_to(np.asarray(sizes), (self.n,))
        return [flat[j, :S * S].reshape(S, S) for j, S in enumerate(sizes)]

    def close(self) -> None:
        if self.id >= 0:
            L.check(L.load().kmc_batch_destroy(self.id))
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MeanField:
    """Result of meanfield_simulate for n networks."""
    occupation: np.ndarray            # (n, NMAX) fractional occupations
    electrode_occupation: np.ndarray  # (n, PMAX) fractional hop counts
    site_energies: np.ndarray         # (n, NMAX) acceptor energies of the last step
    time: np.ndarray                  # (n,)
    average_occupation: np.ndarray    # (n, NMAX) occupation x time, or None
    traffic: np.ndarray               # (n, Smax, Smax) net traffic, or None


def meanfield_simulate(layouts, esets, steps, occupation, record=False) -> MeanField:
    """probSimulate (goSimulation/probabilitySimulation.go:53-157) on the GPU for
    n networks: Layout objects layouts[i] with energy sets esets[i], fractional
    start occupations (n, N_i or NMAX).  Sums run in another order than the
    reference's loops (rounding-level differences)."""
    layouts = list(layouts)
    lay = np.ascontiguousarray([x.id for x in layouts], dtype=np.int32)
    es = np.ascontiguousarray(np.atleast_1d(esets), dtype=np.int32)
    n = lay.size
    occ0 = np.atleast_2d(np.asarray(occupation, dtype=np.float64))
    occ = np.zeros((n, L.NMAX))
    occ[:, :occ0.shape[1]] = occ0
    eo = np.zeros((n, L.PMAX))
    se = np.zeros((n, L.NMAX))
    t = np.zeros(n)
    avg = np.zeros((n, L.NMAX)) if record else None
    smax = max(x.S for x in layouts)
    tr = np.zeros((n, smax, smax)) if record else None
    L.check(L.load().kmc_meanfield_simulate(
        n, L.ptr(lay, L._i32p), L.ptr(es, L._i32p), int(steps), L.ptr(occ, L._dp),
        L.ptr(eo, L._dp), L.ptr(se, L._dp), L.ptr(t, L._dp), L.ptr(avg, L._dp), L.ptr(tr, L._dp)))
    r = MeanField()
    r.occupation, r.electrode_occupation, r.site_energies, r.time = occ, eo, se, t
    r.average_occupation, r.traffic = avg, tr
    return r


def average_current(layouts, esets, occupation) -> np.ndarray:
    """calcAverageCurrent (goSimulation/simulation.go:671-736) on the GPU at the
    time-averaged occupations (n, N_i or NMAX) of Layout objects layouts[i] with
    energy sets esets[i]; returns currents (n, PMAX)."""
    lay = np.ascontiguousarray([x.id for x in layouts], dtype=np.int32)
    es = np.ascontiguousarray(np.atleast_1d(esets), dtype=np.int32)
    n = lay.size
    occ0 = np.atleast_2d(np.asarray(occupation, dtype=np.float64))
    occ = np.zeros((n, L.NMAX))
    occ[:, :occ0.shape[1]] = occ0
    cur = np.zeros((n, L.PMAX))
    L.check(L.load().kmc_average_current(n, L.ptr(lay, L._i32p), L.ptr(es, L._i32p),
                                         L.ptr(occ, L._dp), L.ptr(cur, L._dp)))
    return cur


def simulate_discrete(N, P, nu, kT, I_0, R, distances, E_constant, site_energies,
                      transitions_constant, occupation, electrode_occupation, hops, record=False,
                      uniforms=None, seed=0, hop0=0, log=False):
    """One network through kmc_simulate_discrete (FP64 dense kernel, one launch).

    ``site_energies``, ``occupation`` and ``electrode_occupation`` are updated in
    place like the reference's numba function updates its arguments.  Returns
    (time, traffic, occupations_in_time, events, totals).
    """
    S = int(N) + int(P)
    d = np.ascontiguousarray(distances, dtype=np.float64)
    tc = np.ascontiguousarray(transitions_constant, dtype=np.float64)
    E = np.ascontiguousarray(E_constant, dtype=np.float64)
    se = site_energies
    if se.dtype != np.float64 or not se.flags.c_contiguous:
        raise TypeError("site_energies must be a contiguous float64 array")
    occ = np.ascontiguousarray(occupation, dtype=np.uint8)
    eo = np.ascontiguousarray(electrode_occupation, dtype=np.int64)
    traffic = np.zeros((S, S)) if record else None
    oit = np.zeros(int(N)) if record else None
    u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64).ravel()
    if u is not None and u.size < 2 * hops:
        raise ValueError("need 2 uniforms per hop")
    ev = np.zeros(hops, dtype=np.int32) if log else None
    tot = np.zeros(hops) if log else None
    t = C.c_double()
    rc = L.load().kmc_simulate_discrete(
        int(N), int(P), float(nu), float(kT), float(I_0), float(R), L.ptr(d, L._dp),
        L.ptr(E, L._dp), L.ptr(se, L._dp), L.ptr(tc, L._dp), L.ptr(occ, L._u8p),
        L.ptr(eo, L._i64p), int(hops), int(bool(record)), L.ptr(traffic, L._dp),
        L.ptr(oit, L._dp), L.ptr(u, L._dp), int(seed), int(hop0), L.ptr(ev, L._i32p),
        L.ptr(tot, L._dp), C.byref(t))
    occupation[...] = occ.astype(occupation.dtype)
    electrode_occupation[...] = eo
    L.check(rc)
    return t.value, traffic, oit, ev, tot
