"""Serving loop: continuous batching of KMC evaluation requests on one GPU.

A request names one of the resident layouts, carries one or more electrode
voltage vectors (an IV sweep or a voltage-search candidate) with a number of
replicas, and a hop budget per trajectory; it is the evaluation that
kmc_dn.go_simulation performs per voltage (kmc_dopant_networks.py:473-511,
driven by IV, kmc_dopant_networks_utils.py:639-693, and by the voltage search,
voltage_search.py:102-157).

* Layouts (pair tables and site positions) stay resident; each request's
  energy sets are written into a per-layout pool of device energy sets
  (E_constant from the layout's unit-voltage Laplace basis).
* Admission is FIFO by device memory: a request is admitted when the bytes of
  its trajectories' device state (descriptor, output record, group entry and
  energy set) fit in the budget beside the admitted ones, and free trajectory
  slots remain.
* Work proceeds in rounds (iteration-level scheduling).  Each round runs every
  admitted trajectory over its next hop range in one launch of the batched
  kernel; requests that arrived meanwhile join at the next round, finished ones
  leave.  Ranges start at multiples of the recompute interval, so a chunked
  trajectory continues exactly as one long launch would, and every trajectory's
  Philox key is derived from its request id: a request's result is the same
  alone and in any mix (the pair-term form per size bucket is fixed for the
  server's lifetime).
* A request whose trajectory stops (zero total rate) or whose inputs are
  rejected completes with its status; the others are not affected.
* A request cancelled by the caller (Server.cancel) or past its deadline
  (ServeRequest.deadline_s after its arrival) completes at the next round
  boundary with status ERR_CANCELLED and its tallies so far; its slots and
  energy sets return to the pools.  A watchdog bounds every round's wait: a
  round whose launch has not ended within watchdog_s raises (the device
  stopped making progress; rounds end by themselves at their time limit).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import autotune, engine

TRAJ_DEV_BYTES = 88      # TrajDev
TRAJ_OUT_BYTES = 184     # TrajOut
GROUP_BYTES = 16         # GroupDev


def bytes_per_trajectory(N: int, P: int) -> int:
    """Device bytes of one admitted trajectory's state."""
    return TRAJ_DEV_BYTES + TRAJ_OUT_BYTES + GROUP_BYTES + 8 * (N + P)


def trajectory_seed(rid: int, j: int) -> int:
    """Philox key of trajectory j of request rid."""
    return (int(rid) << 24) | int(j)


def popcount_rows(bits: np.ndarray) -> np.ndarray:
    """Set bits per row of a (n, OCCW) uint64 array."""
    return np.unpackbits(np.ascontiguousarray(bits).view(np.uint8), axis=1).sum(1)


@dataclass
class ServeRequest:
    rid: int
    t_arrive: float
    layout: int                   # index into the server's layouts
    V: np.ndarray                 # (P,) or (k, P) voltage vectors
    hops: int
    replicas: int = 1
    deadline_s: float | None = None   # seconds after t_arrive; None: none


@dataclass
class Completed:
    rid: int
    t_arrive: float
    t_admit: float
    t_done: float
    status: int                   # 0 ok, else the first failing trajectory's code
    currents: np.ndarray          # (k, replicas, P) eo/time per trajectory
    eo: np.ndarray                # (k, replicas, PMAX)
    time: np.ndarray              # (k, replicas)
    occ_bits: np.ndarray          # (k, replicas, OCCW)
    message: str = ""

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrive


class _Active:
    def __init__(self, req: ServeRequest, t_admit: float, slots: np.ndarray, esets: list,
                 nbytes: int):
        self.req = req
        self.t_admit = t_admit
        self.slots = slots
        self.esets = esets
        self.bytes = nbytes
        self.left = slots.size


class Server:
    """Continuous-batching evaluator over resident layouts.

    nets:  workloads.Network per layout; bases: laplace.LaplaceBasis per layout.
    chunk_hops: sets the round's target device time, chunk_hops x 1.5 us; each
      trajectory's hop range per round is that time over its measured seconds per
      hop (a multiple of the recompute interval).
    max_active / mem_budget: trajectory slots and device bytes of admitted state.
    team_min_hops: budgets from which a request runs in a team bucket (None: solo
      buckets only; every admitted trajectory must be resident in each round, and
      table-form teams hold a CU's LDS each).
    plan: fixed launch plan (else autotuned once at start-up).
    """

    REF_COST = 32.0           # cost-model units of a small network (one chunk, 16 walked)
    REF_HOP_S = 1.5e-6        # first guess of the seconds per hop at REF_COST

    def __init__(self, nets, bases, chunk_hops: int = 8192, max_active: int = 8192,
                 mem_budget: int = 64 << 20, plan=None, recompute_interval: int = 1024,
                 eset_block: int = 64, team_min_hops: int | None = None, tune_log=None,
                 watchdog_s: float = 30.0):
        if chunk_hops <= 0 or chunk_hops % recompute_interval:
            raise ValueError("chunk_hops must be a positive multiple of recompute_interval")
        self.nets, self.bases = list(nets), list(bases)
        self.K = recompute_interval
        self.chunk = chunk_hops
        self.round_s = chunk_hops * self.REF_HOP_S      # target device time of a round
        self.cal = self.REF_HOP_S / self.REF_COST       # seconds per hop per cost unit
        self.max_active = max_active
        self.mem_budget = mem_budget
        self.eset_block = eset_block
        self.layouts, self.sizes = [], {}
        for net in self.nets:
            lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                                net.transitions_constant)
            lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
            self.layouts.append(lay)
            self.sizes[lay.id] = (lay.S, lay.N)
        self.lid = np.array([lay.id for lay in self.layouts], np.int32)
        self.lN = np.array([n.N for n in self.nets], np.int64)
        self.lP = np.array([n.P for n in self.nets], np.int64)
        self.lC = (self.lN + self.lP + 63) // 64
        self.occ0 = [engine.occupation_to_bits(n.occupation) for n in self.nets]
        self.occ0_table = np.stack(self.occ0)
        self.host_s = {"admit": 0.0, "prepare": 0.0, "kernel_wait": 0.0, "results": 0.0}
        self.free = collections.defaultdict(list)   # layout index -> free energy-set ids
        # slot table (one row per admitted trajectory)
        n = max_active
        self.s_live = np.zeros(n, bool)
        self.s_req = np.zeros(n, np.int64)
        self.s_lay = np.zeros(n, np.int64)
        self.s_eset = np.zeros(n, np.int32)
        self.s_seed = np.zeros(n, np.uint64)
        self.s_hops = np.zeros(n, np.int64)
        self.s_h = np.zeros(n, np.int64)
        self.s_time = np.zeros(n)
        self.s_eo = np.zeros((n, L.PMAX), np.int64)
        self.s_occ = np.zeros((n, L.OCCW), np.uint64)
        self.s_status = np.zeros(n, np.int32)
        self.s_lat = np.zeros(n)          # seconds per hop: measured, else modelled
        self.free_slots = list(range(n - 1, -1, -1))
        self.team_min_hops = team_min_hops
        self.latency = {}
        self.plan = [tuple(int(x) for x in p) for p in plan] if plan is not None \
            else self._tune(tune_log)
        # rounds end at a device time limit per trajectory; team kernels carry none
        # (their waves would have to agree on where to stop), so plans with team
        # entries size each round by its hop ranges alone
        teams = any(len(e) > 3 and e[3] > 1 for e in self.plan)
        self.time_limit_us = 0 if teams else int(self.round_s * 1e6)

        self.batch = None
        self.pending = collections.deque()
        self.active = {}
        self.mem_used = 0
        self.completed = []
        self.rounds = 0
        self.round_ms = []
        self.round_stats = []     # (trajectories, kernel ms, slowest and median trajectory ms)
        self.hops_done = 0
        self._next = 0
        self.watchdog_s = watchdog_s
        self.cancelled = set()       # request ids to cancel at the next round boundary

    # ---- set-up -------------------------------------------------------------
    def _tune(self, log):
        """Launch plan timed on up to 32 trajectories per layout (at most 2048): solo
        buckets tuned for the throughput of a batch larger than a loaded round, so
        that a candidate whose workgroups per CU are limited by its table's LDS
        pays for the slots it lacks (rounds under load hold more trajectories than
        such a kernel has resident slots); team buckets (budgets >=
        team_min_hops) for hop latency.  A request's bucket depends only on its
        layout and budget."""
        reps = max(1, min(32, 2048 // len(self.nets)))
        t = np.zeros(len(self.nets) * reps, dtype=engine.TRAJ_DTYPE)
        taken = []
        for i, net in enumerate(self.nets):
            es = self._take_eset(i)
            taken.append((i, es))
            sl = slice(i * reps, (i + 1) * reps)
            t["layout"][sl], t["eset"][sl] = self.layouts[i].id, es
            t["seed"][sl] = np.arange(reps)
            t["occ"][sl] = self.occ0[i]
        edges = autotune.BUCKET_EDGES
        if self.team_min_hops is None:   # throughput-tuned solo buckets only
            plan = [tuple(e) for e in autotune.tune(t, self.sizes, log=log)]
        else:
            plan = autotune.plan_for(t, self.sizes, log=log,
                                     team_min_hops={e: self.team_min_hops for e in edges})
        if self.team_min_hops is not None:
            self.latency = autotune.tune_latency(t, self.sizes, log=log)
        for i, es in taken:
            self._give_eset(i, es)
        return plan

    def _take_eset(self, li: int) -> int:
        if not self.free[li]:
            net = self.nets[li]
            E = np.repeat(net.E_constant[None], self.eset_block, 0)
            V = np.repeat(net.site_energies[None, net.N:], self.eset_block, 0)
            self.free[li].extend(int(x) for x in self.layouts[li].energy_sets(E, V)[::-1])
        return self.free[li].pop()

    def _give_eset(self, li: int, eset: int) -> None:
        self.free[li].append(eset)

    def energy(self, li: int, V: np.ndarray) -> np.ndarray:
        """E_constant of layout li at electrode voltages V (P,)."""
        return self.bases[li].eV(V) + self.nets[li].comp_constant

    # ---- requests -----------------------------------------------------------
    def submit(self, req: ServeRequest) -> None:
        self.pending.append(req)

    def busy(self) -> bool:
        return bool(self.pending or self.active)

    def cancel(self, rid: int) -> None:
        """Cancel request `rid` (pending or running) at the next round boundary."""
        self.cancelled.add(int(rid))

    def _late(self, req, now: float) -> bool:
        return int(req.rid) in self.cancelled or (
            req.deadline_s is not None and now - req.t_arrive > req.deadline_s)

    def _expire(self, now: float) -> None:
        """Complete cancelled requests and those past their deadline: pending ones
        without running, running ones with their tallies so far."""
        if not self.cancelled and not any(r.deadline_s is not None for r in self.pending) and \
                not any(a.req.deadline_s is not None for a in self.active.values()):
            return
        keep = collections.deque()
        for req in self.pending:
            if self._late(req, now):
                self._fail(req, np.atleast_2d(np.asarray(req.V, dtype=np.float64)), now,
                           L.ERR_CANCELLED, f"request {req.rid} cancelled before it ran")
                self.cancelled.discard(int(req.rid))
            else:
                keep.append(req)
        self.pending = keep
        for rq in [k for k, a in self.active.items() if self._late(a.req, now)]:
            act = self.active[rq]
            sl = act.slots
            ran, total = int(self.s_h[sl].sum()), int(self.s_hops[sl].sum())
            self._complete(rq, now, L.ERR_CANCELLED,
                           f"request {act.req.rid} cancelled after {ran} of {total} hops")
            self.cancelled.discard(int(act.req.rid))

    def _admit(self, now: float) -> None:
        """Admit the longest FIFO prefix of pending requests that fits, in bulk: one
        energy-set upload and vectorised slot assignment per round."""
        t0 = time.perf_counter()
        cand = []
        free, mem = len(self.free_slots), self.mem_used
        while self.pending:
            req = self.pending[0]
            V = np.atleast_2d(np.asarray(req.V, dtype=np.float64))
            if not 0 <= req.layout < len(self.nets):
                self.pending.popleft()
                self._fail(req, V, now, L.ERR_ARGS, f"request {req.rid}: no layout {req.layout}")
                continue
            net = self.nets[req.layout]
            if V.shape[1] != net.P or not np.isfinite(V).all() or V.shape[0] == 0:
                self.pending.popleft()
                self._fail(req, V, now, L.ERR_ARGS,
                           f"request {req.rid}: need finite voltages of shape (k, {net.P})")
                continue
            rep = max(1, int(req.replicas))
            ntask = V.shape[0] * rep
            need = ntask * bytes_per_trajectory(net.N, net.P)
            if ntask > free or mem + need > self.mem_budget:
                if not self.active and not cand:   # could never fit: fail, do not block
                    self.pending.popleft()
                    self._fail(req, V, now, L.ERR_LIMIT, f"request {req.rid} exceeds the limits")
                    continue
                break
            self.pending.popleft()
            req.V = V
            cand.append((req, rep, ntask, need))
            free -= ntask
            mem += need
        if cand:
            self._admit_bulk(cand, now)
        self.host_s["admit"] += time.perf_counter() - t0

    def _admit_bulk(self, cand, now: float) -> None:
        nsets = sum(c[0].V.shape[0] for c in cand)
        E_pad = np.zeros((nsets, L.NMAX))
        V_pad = np.zeros((nsets, L.PMAX))
        ids = np.zeros(nsets, np.int32)
        set_lay = np.zeros(nsets, np.int64)
        row = 0
        per_req_sets = []
        for req, _, _, _ in cand:
            k, P = req.V.shape[0], self.nets[req.layout].P
            es = [self._take_eset(req.layout) for _ in range(k)]
            ids[row:row + k] = es
            set_lay[row:row + k] = req.layout
            V_pad[row:row + k, :P] = req.V
            per_req_sets.append(es)
            row += k
        for li in np.unique(set_lay):
            rows = np.nonzero(set_lay == li)[0]
            net = self.nets[li]
            E_pad[rows, :net.N] = self.bases[li].eV(V_pad[rows, :net.P]) + net.comp_constant
        engine.Layout.write_energy_sets(ids, E_pad, V_pad)
        ntasks = np.array([c[2] for c in cand], np.int64)
        total = int(ntasks.sum())
        slots = np.array([self.free_slots.pop() for _ in range(total)], np.int64)
        first = np.concatenate([[0], np.cumsum(ntasks)[:-1]])
        j = np.arange(total) - np.repeat(first, ntasks)          # index inside its request
        rq = np.repeat(np.arange(self._next, self._next + len(cand)), ntasks)
        rid = np.repeat(np.array([c[0].rid for c in cand], np.uint64), ntasks)
        lay = np.repeat(np.array([c[0].layout for c in cand], np.int64), ntasks)
        hops = np.repeat(np.array([max(0, int(c[0].hops)) for c in cand], np.int64), ntasks)
        es = np.concatenate([np.repeat(sets, c[1]) for sets, c in zip(per_req_sets, cand)])
        self.s_live[slots] = True
        self.s_req[slots] = rq
        self.s_lay[slots] = lay
        self.s_eset[slots] = es
        self.s_seed[slots] = (rid << np.uint64(24)) | j.astype(np.uint64)
        self.s_hops[slots] = hops
        self.s_h[slots] = 0
        self.s_time[slots] = 0.0
        self.s_eo[slots] = 0
        self.s_occ[slots] = self.occ0_table[lay]
        self.s_status[slots] = 0
        self.s_lat[slots] = self.cost(slots) * self.cal
        off = 0
        for (req, _, ntask, need), sets in zip(cand, per_req_sets):
            self.active[self._next] = _Active(req, now, slots[off:off + ntask], sets, need)
            self.mem_used += need
            self._next += 1
            off += ntask

    def _fail(self, req, V, now, code, msg) -> None:
        k, r = V.shape[0], max(1, int(req.replicas))
        P = self.nets[req.layout].P if 0 <= req.layout < len(self.nets) else 0
        self.completed.append(Completed(req.rid, req.t_arrive, now, now, int(code),
                                        np.full((k, r, P), np.nan),
                                        np.zeros((k, r, L.PMAX), np.int64), np.zeros((k, r)),
                                        np.zeros((k, r, L.OCCW), np.uint64), msg))

    # ---- one round ------------------------------------------------------------
    def _ensure_batch(self) -> None:
        if self.batch is not None:
            return
        t = np.zeros(self.max_active, dtype=engine.TRAJ_DTYPE)
        es = self._take_eset(0)
        self._give_eset(0, es)
        t["layout"], t["eset"] = self.layouts[0].id, es
        self.batch = engine.Batch(t)
        self.batch.set_buckets(self.plan)

    def cost(self, idx: np.ndarray) -> np.ndarray:
        """Static per-hop cost model, relative units: site chunks x (walked-set size
        + fixed overhead), divided by the hop-latency gain of a team bucket."""
        li = self.s_lay[idx]
        N, P, C = self.lN[li], self.lP[li], self.lC[li]
        nocc = popcount_rows(self.s_occ[idx])
        walk = np.minimum(nocc + P, N - nocc + P)
        gain = self.team_gain(idx) if getattr(self, "latency", None) else 1.0
        return C * (16 + walk) / gain

    def quotas(self, idx: np.ndarray) -> np.ndarray:
        """Hop range per trajectory this round: twice the round's target time over
        the trajectory's seconds per hop (measured in its previous round, else the
        cost model times the current calibration), in multiples of the recompute
        interval.  The device time limit, not the range, ends a round: a hop's
        latency varies several-fold along a trajectory (runs of hops between two
        cached states against hops into new states), and one trajectory whose
        range was sized on a fast round held a whole round up to ten times its
        target; the range only keeps a trajectory from running past a round far
        beyond its measured pace.  Without a time limit (plans with team entries)
        the range is the round's target time over that pace."""
        lat = self.s_lat[idx]
        if not getattr(self, "time_limit_us", 0):   # the ranges alone size the round
            q = (self.round_s / lat).astype(np.int64) // self.K * self.K
            return np.clip(q, self.K, 16 * self.chunk)
        q = (2.0 * self.round_s / lat).astype(np.int64) // self.K * self.K
        return np.clip(q, self.K, 64 * self.chunk)

    def plan_entry(self, S: np.ndarray, budget: np.ndarray) -> np.ndarray:
        """Index of the plan entry each trajectory runs in: the first entry with
        S <= s_max and budget >= min_hops (the engine's rule)."""
        out = np.full(S.shape, -1)
        for k in range(len(self.plan) - 1, -1, -1):
            e = self.plan[k]
            out[(S <= e[0]) & (budget >= (e[4] if len(e) > 4 else 0))] = k
        return out

    def team_gain(self, idx: np.ndarray) -> np.ndarray:
        """Hop-latency gain (solo / team) of the plan entry each trajectory runs in."""
        gain = np.ones(idx.size)
        S = self.lN[self.s_lay[idx]] + self.lP[self.s_lay[idx]]
        entry = self.plan_entry(S, self.s_hops[idx])
        for k, e in enumerate(self.plan):
            lt = self.latency.get(e[0])
            if len(e) > 3 and e[3] > 1 and lt:
                solo = min(lt["solo"].values()) if lt["solo"] else lt["lat"]
                gain[entry == k] = max(1.0, solo / lt["lat"])
        return gain

    def round(self, clock) -> int:
        """Admit, run one round, retire finished requests; returns trajectories run."""
        self._expire(clock())
        self._admit(clock())
        idx = np.nonzero(self.s_live & (self.s_h < self.s_hops))[0]
        self._retire(clock)
        if idx.size == 0:
            return 0
        self._ensure_batch()
        t0 = time.perf_counter()
        n = idx.size
        trajs = np.zeros(n, dtype=engine.TRAJ_DTYPE)
        trajs["layout"] = self.lid[self.s_lay[idx]]
        trajs["eset"] = self.s_eset[idx]
        trajs["seed"] = self.s_seed[idx]
        trajs["hops"] = self.s_hops[idx]
        trajs["occ"] = self.s_occ[idx]
        hb = self.s_h[idx]
        he = np.minimum(self.s_hops[idx], hb + self.quotas(idx))
        b = self.batch
        b.update(trajs)
        b.set_state(hb, he, self.s_time[idx], self.s_eo[idx], self.s_occ[idx])
        # the device time limit ends the round: a trajectory stops at its first
        # recompute point past it and continues there next round
        b.run(precision=L.FP32_BATCHED, resume=True, recompute_interval=self.K,
              time_limit_us=self.time_limit_us)
        t1 = time.perf_counter()
        ms = b.wait(self.watchdog_s * 1e3)
        if ms is None:
            raise RuntimeError(f"serving round {self.rounds} did not end within "
                               f"{self.watchdog_s} s ({n} trajectories)")
        self.round_ms.append(ms)
        t2 = time.perf_counter()
        r = b.results()
        # hops reached: the range's end, or the recompute point where the time
        # limit stopped the trajectory (the budget for one that stopped early)
        reached = np.where(r.status != 0, self.s_hops[idx], r.hops_done)
        ran = reached - hb
        sec = b.timing()
        m = (ran >= 256) & (sec > 0)
        if m.any():
            lat = sec[m] / ran[m]
            self.s_lat[idx[m]] = lat
            self.cal = float(np.median(lat / self.cost(idx[m])))
        self.round_stats.append((n, self.round_ms[-1], float(sec.max()) * 1e3,
                                 float(np.median(sec)) * 1e3))
        self.s_time[idx] = r.time
        self.s_eo[idx] = r.eo
        self.s_occ[idx] = r.occ_bits
        self.s_status[idx] = r.status
        self.s_h[idx] = reached
        self.hops_done += int(np.sum(ran))
        self.rounds += 1
        self._retire(clock)
        t3 = time.perf_counter()
        self.host_s["prepare"] += t1 - t0
        self.host_s["kernel_wait"] += t2 - t1
        self.host_s["results"] += t3 - t2
        return n

    def _retire(self, clock) -> None:
        """Completes every active request whose trajectories all reached their budget."""
        done = self.s_live & (self.s_h >= self.s_hops)
        if not done.any():
            return
        now = clock()
        for rq in np.unique(self.s_req[done]):
            act = self.active[int(rq)]
            sl = act.slots
            if not (self.s_h[sl] >= self.s_hops[sl]).all():
                continue
            st = self.s_status[sl]
            self._complete(int(rq), now, int(st[st != 0][0]) if (st != 0).any() else 0)

    def _complete(self, rq: int, now: float, status: int, message: str = "") -> None:
        """Completes active request rq with its trajectories' current tallies and
        returns its slots and energy sets."""
        act = self.active[rq]
        sl = act.slots
        req = act.req
        k, rep = req.V.shape[0], max(1, int(req.replicas))
        P = self.nets[req.layout].P
        eo = self.s_eo[sl].reshape(k, rep, L.PMAX).copy()
        tm = self.s_time[sl].reshape(k, rep).copy()
        with np.errstate(divide="ignore", invalid="ignore"):
            cur = eo[:, :, :P] / tm[:, :, None]
        self.completed.append(Completed(req.rid, req.t_arrive, act.t_admit, now, status, cur,
                                        eo, tm, self.s_occ[sl].reshape(k, rep, L.OCCW).copy(),
                                        message))
        self.s_live[sl] = False
        self.free_slots.extend(int(x) for x in sl)
        for es in act.esets:
            self._give_eset(req.layout, es)
        self.mem_used -= act.bytes
        del self.active[rq]

    def close(self) -> None:
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for lay in self.layouts:
            lay.close()
        self.layouts = []


def serve_stream(server: Server, requests, realtime: bool = True, max_wall: float = 120.0):
    """Feed `requests` (sorted by t_arrive, seconds from start) into the server in
    real time and run rounds until all complete; returns (completed, wall seconds)."""
    t0 = time.perf_counter()

    def clock():
        return time.perf_counter() - t0

    reqs = list(requests)
    i = 0
    first = len(server.completed)
    while i < len(reqs) or server.busy():
        if clock() > max_wall:
            raise TimeoutError(f"stream not drained after {max_wall} s")
        now = clock()
        while i < len(reqs) and (not realtime or reqs[i].t_arrive <= now):
            server.submit(reqs[i])
            i += 1
        if server.round(clock) == 0 and not server.busy() and i < len(reqs) and realtime:
            time.sleep(max(0.0, min(reqs[i].t_arrive - clock(), 0.05)))
    return server.completed[first:], clock()


def serve_one_launch_per_request(server: Server, requests, realtime: bool = True,
                                 max_requests: int | None = None):
    """Baseline: each request runs alone, in one launch over its whole budget, as
    soon as it has arrived and the previous launch has finished."""
    reqs = list(requests)[:max_requests]
    t0 = time.perf_counter()
    out = []
    batch = None
    try:
        for req in reqs:
            now = time.perf_counter() - t0
            if realtime and req.t_arrive > now:
                time.sleep(req.t_arrive - now)
            t_start = time.perf_counter() - t0
            li = req.layout
            net = server.nets[li]
            V = np.atleast_2d(np.asarray(req.V, dtype=np.float64))
            k, rep = V.shape[0], max(1, int(req.replicas))
            esets = []
            for kk in range(k):
                es = server._take_eset(li)
                esets.append(es)
                server.layouts[li].write_energy_set(es, server.energy(li, V[kk]), V[kk])
            t = np.zeros(k * rep, dtype=engine.TRAJ_DTYPE)
            t["layout"] = server.layouts[li].id
            t["eset"] = np.repeat(esets, rep)
            t["seed"] = (np.uint64(req.rid) << np.uint64(24)) | np.arange(k * rep, dtype=np.uint64)
            t["hops"] = req.hops
            t["occ"] = server.occ0[li]
            if batch is None:
                batch = engine.Batch(t)
                batch.set_buckets(server.plan)
            else:
                batch.update(t)
            batch.run(precision=L.FP32_BATCHED, recompute_interval=server.K)
            batch.sync()
            r = batch.results()
            t_done = time.perf_counter() - t0
            for es in esets:
                server._give_eset(li, es)
            P = net.P
            eo = r.eo.reshape(k, rep, L.PMAX)
            tm = r.time.reshape(k, rep)
            with np.errstate(divide="ignore", invalid="ignore"):
                cur = eo[:, :, :P] / tm[:, :, None]
            st = int(r.status[r.status != 0][0]) if (r.status != 0).any() else 0
            out.append(Completed(req.rid, req.t_arrive, t_start, t_done, st, cur, eo, tm,
                                 r.occ_bits.reshape(k, rep, L.OCCW)))
    finally:
        if batch is not None:
            batch.close()
    return out, time.perf_counter() - t0


def latency_summary(done, wall: float, hops: int | None = None) -> dict:
    lat = np.array([c.latency for c in done if c.status == 0])
    d = {"completed": len(done), "failed": int(sum(c.status != 0 for c in done)),
         "wall_s": wall, "req_per_s": len(done) / wall if wall > 0 else None}
    if lat.size:
        d.update(p50_ms=float(np.percentile(lat, 50) * 1e3),
                 p99_ms=float(np.percentile(lat, 99) * 1e3),
                 mean_ms=float(lat.mean() * 1e3), max_ms=float(lat.max() * 1e3))
    if hops is not None:
        d["hops_per_s"] = hops / wall if wall > 0 else None
    return d
