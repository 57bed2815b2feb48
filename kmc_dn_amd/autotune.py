"""Autotuner of the FP32 batched kernel's launch shape per network-size bucket.

For every size bucket (S <= 64, <= 128, <= 192: one, two or three 64-site
chunks per lane) it times the candidate pair-term forms

  1  {lt, lt+c} pair table + c table in LDS, shared by the W waves of a workgroup
  2  {lt, c} pair table in LDS
  3  site positions in LDS, pair terms computed (no table)

times the candidate workgroup sizes W in {1, 2, 4, 8} on a sample of the
bucket's own trajectories (their layouts, energy sets and seeds, a short hop
budget), and keeps the fastest.  Forms whose LDS footprint does not fit 160 KiB
are skipped.  Results are cached in a JSON file keyed by device, bucket and the
trajectories-per-layout class of the sample, so later runs of a similar batch
reuse them without timing.

No reference counterpart: the reference runs one network per call
(goSimulation/simulationWrapper.go:83-169, kmc_dopant_networks.py:33-135).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import _lib as L
from . import engine

BUCKET_EDGES = (64, 128, 192)
FORMS = (1, 2, 3)
WAVES = (1, 2, 4, 8)
LDS_BYTES = 160 * 1024
# concurrent bucket launches: HIP gives a process 4 hardware queues by default and
# the default stream (host copies) holds one of them
MAX_LAUNCHES = 3
# hop latency of a solo trajectory inside a loaded batch over its latency alone:
# in the 2048-network uneven batch with the per-trajectory state cache, an
# S = 130 trajectory in the positions-form solo entry ran at 4.19 us/hop against
# 1.33 alone (3.15x), and under a table-form solo-only plan the longest
# trajectories ran at 2.0-2.7 us/hop (S = 66-107) against 1.1 alone: a solo
# wave shares its SIMD with other trajectories' waves, a team's waves do not
BATCH_SLOWDOWN = 3.0
# least ratio of the solo latency inside the batch to the team latency alone for
# a team entry
TEAM_MIN_GAIN = 1.2
# team size of team entries: four waves fill a CU's four SIMDs, and a table
# large enough to need two or more site chunks keeps other workgroups off most
# of the CU's LDS.  With the state cache, teams of 1, 2 and 4 waves are within
# 5 % of one another alone (0.96-0.99 us/hop at S = 132); in the batch, teams of
# 4 for S > 64 gave 1.15 s and an autotuned plan of teams of 2 with a
# one-chunk team entry 1.89 s.
TEAM_SIZE = 4
TEAM_MIN_EDGE = 128
# Latency probe of the team thresholds (probe_team_thresholds): warm-up hops,
# timed hops, and the ratios measured on the 2048-network batch: a team's hop
# latency over the probed one (S = 81: 1.05 against 2.26 us/hop), a solo
# trajectory's inside the batch over the probed one
PROBE_WARM = 4096
PROBE_HOPS = 16384
PROBE_TEAM = 0.5
PROBE_SOLO = 1.5
CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")


def table_lds_bytes(form: int, S: int, N: int, waves: int = 8) -> int:
    """Upper bound of the LDS bytes of a table-form workgroup of `waves` waves
    (or a team of as many) for layouts of at most S sites and N acceptors: the
    waves' electrode factors, the electrode weights of up to 16 electrodes, the
    pair table and, for form 1, the c table (kmc_batch.hip batch_lds_bytes)."""
    C = (S + 63) // 64
    b = waves * 2 * 16 * 4 + 16 * 64 * C * 4 + (S * S + 64 * C) * 8
    if form == 1:
        b += (N * S + 64 * C) * 4
    return b


def bucket_of(S: int, edges=BUCKET_EDGES) -> int:
    for b, e in enumerate(edges):
        if S <= e:
            return b
    raise ValueError(f"S={S} above the largest bucket {edges[-1]}")


def candidates(smax: int, nmax: int, has_pos: bool):
    for form in FORMS:
        if form == 3 and not has_pos:
            continue
        for w in WAVES:
            if form != 3 and table_lds_bytes(form, smax, nmax, w) > LDS_BYTES:
                continue
            yield form, w


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(path: str, cache: dict) -> None:
    tmp = path + ".tmp"
    try:
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass   # a read-only tree just skips caching


def cache_key(device: str, edge: int, traj_per_layout: float) -> str:
    return f"{device}|S<={edge}|tpl{int(round(math.log2(max(1.0, traj_per_layout))))}"


def time_candidate(trajs: np.ndarray, edge: int, form: int, w: int, reps: int = 2) -> float:
    """Hops per second of one candidate on `trajs` (best of `reps` timed runs)."""
    b = engine.Batch(trajs)
    try:
        b.set_buckets([(edge, form, w, 0)])
        b.run(precision=L.FP32_BATCHED)
        b.sync()
        best = math.inf
        for _ in range(reps):
            b.run(precision=L.FP32_BATCHED)
            best = min(best, b.sync())
        r = b.results()
        if (r.status != 0).any():
            raise L.KmcError(L.ERR_ZERO_RATE, "a tuning trajectory stopped")
        hops = float(np.sum(trajs["hops"] + trajs["prehops"]))
        return hops / (best * 1e-3)
    finally:
        b.close()


def time_latency(traj: np.ndarray, form: int, team: int, hops: int = 4096) -> float:
    """Seconds per hop of one trajectory running alone (the hop latency)."""
    t = traj.copy().reshape(1)
    t["prehops"] = 0
    t["hops"] = hops
    b = engine.Batch(t)
    try:
        b.run(precision=L.FP32_BATCHED, table_form=form, team=team, waves_per_group=1)
        b.sync()
        b.run(precision=L.FP32_BATCHED, table_form=form, team=team, waves_per_group=1)
        return b.sync() * 1e-3 / hops
    finally:
        b.close()


def tune_latency(trajs: np.ndarray, sizes: dict, has_pos: bool = True, edges=BUCKET_EDGES,
                 cache_path: str | None = CACHE_PATH, retune: bool = False, log=None) -> dict:
    """Per size bucket: the fastest (form, team) for one trajectory alone and its
    seconds per hop, plus the solo latency of every fitting form, measured on the
    bucket's largest network.  Returns {edge: {"form", "team", "lat", "solo": {form: lat}}}."""
    device = engine.device_info()["name"]
    cache = _load_cache(cache_path) if cache_path else {}
    lay = trajs["layout"]
    S_of = np.array([sizes[int(x)][0] for x in lay])
    N_of = np.array([sizes[int(x)][1] for x in lay])
    out, changed = {}, False
    for b, edge in enumerate(edges):
        idx = np.nonzero(np.array([bucket_of(int(x), edges) for x in S_of]) == b)[0]
        if idx.size == 0:
            continue
        big = idx[np.argmax(S_of[idx])]
        key = f"{device}|S<={edge}|latency|S{int(S_of[big])}"
        if key in cache and not retune:
            out[edge] = cache[key]
            continue
        res = {}
        for form in FORMS:
            if form == 3 and not has_pos:
                continue
            if form != 3 and table_lds_bytes(form, int(S_of[big]), int(N_of[big])) + \
                    2 * 4 * 3 * 64 * 4 > LDS_BYTES:
                continue
            for team in (1, 2, 4):
                res[(form, team)] = time_latency(trajs[big], form, team)
                if log:
                    log(f"  latency S={int(S_of[big])}: form {form} team {team}: "
                        f"{res[(form, team)] * 1e6:.2f} us/hop")
        (form, team), lat = min(res.items(), key=lambda kv: kv[1])
        entry = {"form": form, "team": team, "lat": lat,
                 "solo": {str(f): v for (f, t), v in res.items() if t == 1},
                 "teams": {f"{f}/{t}": v for (f, t), v in res.items() if t > 1}}
        cache[key] = entry
        out[edge] = entry
        changed = True
    if changed and cache_path:
        _save_cache(cache_path, cache)
    return out


def plan_for(trajs: np.ndarray, sizes: dict, has_pos: bool = True, edges=BUCKET_EDGES,
             cache_path: str | None = CACHE_PATH, team_min_hops: dict | None = None,
             log=None) -> list:
    """Launch plan for a batch: per size bucket a solo bucket tuned for throughput
    and, in front of it, a team bucket tuned for hop latency that takes the
    trajectories whose solo run would outlast the batch's throughput-bound time;
    merged to the launch limit, then two team entries re-cut by LDS
    (split_team_entries).

    team_min_hops: {edge: budget} to fix the team threshold per bucket instead
    (serving: a property of the request, not of the current mix)."""
    solo = tune(trajs, sizes, has_pos=has_pos, edges=edges, cache_path=cache_path, log=log)
    lat = tune_latency(trajs, sizes, has_pos=has_pos, edges=edges, cache_path=cache_path, log=log)
    cache = _load_cache(cache_path) if cache_path else {}
    device = engine.device_info()["name"]
    lay = trajs["layout"]
    S_of = np.array([sizes[int(x)][0] for x in lay])
    budget = (trajs["hops"] + trajs["prehops"]).astype(np.float64)
    bucket = np.array([bucket_of(int(s), edges) for s in S_of])
    # throughput-bound time: the buckets' work at their tuned rates, run side by side
    t_thr = 0.0
    for edge, form, w, _ in solo:
        b = edges.index(edge)
        sel = bucket == b
        tpl = sel.sum() / max(1, len(np.unique(lay[sel])))
        rate = cache.get(cache_key(device, edge, tpl), {}).get("hops_per_s")
        if rate:
            t_thr = max(t_thr, float(budget[sel].sum()) / rate)
    # team thresholds from the trajectories' own hop latencies (a probe launch);
    # fixed thresholds (serving) and batches without positions keep the model
    probed = None
    if team_min_hops is None and has_pos:
        probed = probe_team_thresholds(trajs, bucket, budget, t_thr, edges)
        if log:
            log(f"  team thresholds from the latency probe (t_thr {t_thr:.3f} s): {probed}")
    teams, solos = [], []
    for edge, form, w, _ in solo:
        b = edges.index(edge)
        lt = lat.get(edge)
        if probed is not None:
            thr = probed.get(edge)
            entry = None if thr is None else team_entry(edge, lt, form, t_thr, budget[bucket == b],
                                                        thr, fixed=True, probed=True)
        else:
            entry = team_entry(edge, lt, form, t_thr, budget[bucket == b],
                               None if team_min_hops is None else team_min_hops.get(edge),
                               fixed=team_min_hops is not None)
        if log and lt:
            log(f"  bucket S<={edge}: team entry {entry} (latency tuning {lt})")
        if entry is not None:
            teams.append(entry)
        solos.append((edge, form, w, 1, 0))
    N_of = np.array([sizes[int(x)][1] for x in lay])
    max_sn = {e: (int(S_of[bucket == b].max()), int(N_of[bucket == b].max()))
              for b, e in enumerate(edges) if (bucket == b).any()}
    return split_team_entries(merge_plan(teams, solos, max_sn, has_pos),
                              int(S_of.max()), int(N_of.max()))


def probe_team_thresholds(trajs: np.ndarray, bucket: np.ndarray, budget: np.ndarray,
                          t_thr: float, edges=BUCKET_EDGES) -> dict:
    """Per size bucket of two or more site chunks: the team threshold from the
    trajectories' own hop latencies, or no entry.

    Hop latencies of networks of one size differ several-fold with how often
    their hops leave the two cached states (1.05 against 0.33 us/hop alone at
    S = 81 and S = 100 in the 2048-network batch), so one network's latency per
    bucket misplaces the long trajectories of the others.  A probe launch runs
    every such trajectory at once (one wave each, positions form) over
    PROBE_WARM hops and then times PROBE_HOPS more (per-trajectory device
    clocks of a resumed range: the state has left its start).  With lat_i that
    time per hop, the batch cannot end before the longest team trajectory,
    PROBE_TEAM * max(budget_i * lat_i); a trajectory goes to the team entry when
    its solo run, PROBE_SOLO * budget_i * lat_i, would end after that or after
    the throughput-bound time t_thr.  The threshold is the smallest budget of
    such a trajectory (entries route by budget)."""
    out = {}
    team_b = [b for b, e in enumerate(edges) if e >= TEAM_MIN_EDGE]
    idx = np.nonzero(np.isin(bucket, team_b))[0]
    if idx.size == 0:
        return out
    n = idx.size
    t = trajs[idx].copy()
    t["prehops"] = 0
    t["hops"] = PROBE_WARM + PROBE_HOPS
    b = engine.Batch(t)
    try:
        b.set_buckets([(edges[-1], 3, 1, 1, 0)])
        b.set_state(h_begin=np.zeros(n, np.int64), h_end=np.full(n, PROBE_WARM, np.int64))
        b.run(precision=L.FP32_BATCHED, resume=True)
        b.sync()
        b.set_state(h_begin=np.full(n, PROBE_WARM, np.int64),
                    h_end=np.full(n, PROBE_WARM + PROBE_HOPS, np.int64))
        b.run(precision=L.FP32_BATCHED, resume=True)
        b.sync()
        r = b.results()
        lat = b.timing() / PROBE_HOPS
    finally:
        b.close()
    if (r.status != 0).any():
        raise L.KmcError(L.ERR_ZERO_RATE, "a probed trajectory stopped")
    work = budget[idx] * lat
    target = max(t_thr, PROBE_TEAM * float(work.max()))
    for bb in team_b:
        m = (bucket[idx] == bb) & (PROBE_SOLO * work > target)
        if m.any():
            out[edges[bb]] = int(budget[idx][m].min())
    return out


def team_lds_bytes(form: int, S: int, N: int, team: int) -> int:
    """Upper bound of the LDS bytes of a table-form team workgroup: the table
    terms plus the team's partial row sums (kmc_team.hip team_lds_bytes)."""
    C = (S + 63) // 64
    return table_lds_bytes(form, S, N, team) + 2 * team * C * 64 * 4


def split_team_entries(plan: list, s_max: int, n_max: int) -> list:
    """Two table-form team entries of one form and team size become one entry
    for the sizes whose team workgroup fits twice into a CU's LDS (S <= cut) and
    one, bounded below by s_min = cut + 1, for the larger sizes, both at the
    lower of the two thresholds.  A team entry whose table takes most of the
    LDS runs one trajectory per CU, so its throughput, not the hop latency of
    its longest trajectory, bounded the 2048-network uneven batch (2.07 s with
    entries cut at S = 128, 1.92-1.94 s cut at S = 90).  The larger entry keeps
    a kernel of as many site chunks as its largest network needs; its table
    must fit the LDS.  Other plans are returned unchanged."""
    teams = [e for e in plan if e[3] > 1]
    if len(teams) != 2 or any(len(e) > 5 for e in teams):
        return plan
    form, team = teams[0][1], teams[0][3]
    if form == 3 or teams[1][1] != form or teams[1][3] != team:
        return plan
    cut = 0
    for S in range(1, s_max):
        if team_lds_bytes(form, S, min(n_max, S), team) <= LDS_BYTES // 2:
            cut = S
    if cut == 0 or team_lds_bytes(form, s_max, n_max, team) > LDS_BYTES:
        return plan
    thr = min(teams[0][4], teams[1][4])
    top = max(teams[0][0], teams[1][0])
    # the smaller entry starts above the one-chunk bucket: one-chunk networks
    # run their long trajectories faster in the solo entry's one-chunk code
    # (fast hops and leaps) than in a team kernel over two chunks (the
    # 2048-network batch: 940 against 967-979 ms)
    low = min(teams[0][0], teams[1][0])
    s_min = max([e for e in BUCKET_EDGES if e < low], default=0) + 1
    return [(top, form, 1, team, thr, 0, cut + 1), (cut, form, 1, team, thr, 0, s_min)] + \
        [e for e in plan if e[3] <= 1]


def team_entry(edge: int, lt: dict | None, solo_form: int, t_thr: float, budgets: np.ndarray,
               min_hops: int | None = None, fixed: bool = False, probed: bool = False):
    """The team entry of one size bucket, or None.

    lt: the bucket's latency tuning ({"form", "team", "lat", "solo": {form: lat},
    "teams": {"form/team": lat}}, seconds per hop of one trajectory alone).  Only
    buckets of two or more site chunks get a team entry, of TEAM_SIZE waves in
    the form fastest alone at that size.  It exists only where that team alone
    is at least TEAM_MIN_GAIN times faster than the best solo form
    inside the loaded batch (BATCH_SLOWDOWN times its lone latency): a marginal
    gain does not pay for the launch slot it takes, and a fourth entry forces
    the larger sizes' team entries to merge into one kernel sized for the
    largest table.  Its threshold takes the trajectories whose solo run (the
    bucket's throughput form, at BATCH_SLOWDOWN times its lone latency) would
    outlast the batch's throughput-bound time t_thr, and only when some budget
    of the bucket reaches it.  fixed: min_hops is the threshold (serving: a
    property of the request, not of the mix).  probed: min_hops came from the
    trajectories' own latencies (probe_team_thresholds), which decided already
    that a team pays; the single network's gain test is skipped."""
    if not lt or not lt["solo"] or edge < TEAM_MIN_EDGE:
        return None
    teams = {k: v for k, v in lt.get("teams", {}).items()}
    if lt["team"] > 1:
        teams.setdefault(f'{lt["form"]}/{lt["team"]}', lt["lat"])
    sized = {k: v for k, v in teams.items() if int(k.split("/")[1]) == TEAM_SIZE}
    teams = sized or teams
    if not teams:
        return None
    tkey, tlat = min(teams.items(), key=lambda kv: kv[1])
    if not probed and tlat * TEAM_MIN_GAIN > min(lt["solo"].values()) * BATCH_SLOWDOWN:
        return None
    if fixed:
        thr = min_hops
    else:
        solo_lat = lt["solo"].get(str(solo_form), lt["lat"]) * BATCH_SLOWDOWN
        thr = int(math.ceil(t_thr / solo_lat)) if t_thr > 0 else None
        if thr is not None and not (np.asarray(budgets) >= thr).any():
            thr = None
    if thr is None:
        return None
    form, team = (int(x) for x in tkey.split("/"))
    return (edge, form, 1, team, int(thr))


def merge_plan(teams: list, solos: list, max_sn: dict, has_pos: bool = True,
               max_launches: int = MAX_LAUNCHES) -> list:
    """Team entries (by size) then solo entries (by size); while there are more
    than `max_launches`, the two largest-size solo entries merge, else the two
    largest-size team entries (while the merged table still fits LDS).  A merged
    entry keeps the larger size's kernel (its site-chunk count covers both
    sizes) and the smaller team threshold; a merged solo entry uses the
    positions form (per-wave coordinates, no per-layout table), which serves
    every size.  Small-size team entries stay apart longest: they hold the long
    trajectories whose hop latency the larger kernels would raise.  Each entry
    is one concurrent launch: HIP runs launches whose streams share a hardware
    queue one after the other."""
    teams, solos = sorted(teams), sorted(solos)
    sn = dict(max_sn)
    red = 2 * 4 * 3 * 64 * 4

    def combined(a, b):
        S = max(sn.get(a[0], (0, 0))[0], sn.get(b[0], (0, 0))[0])
        N = max(sn.get(a[0], (0, 0))[1], sn.get(b[0], (0, 0))[1])
        return S, N

    while len(teams) + len(solos) > max_launches:
        merged = None
        if len(solos) >= 2:
            a, b = solos[-2], solos[-1]
            S, N = combined(a, b)
            if has_pos:
                merged = (solos, (b[0], 3, 4, 1, 0), S, N)
            elif table_lds_bytes(b[1], S, N) <= LDS_BYTES:
                merged = (solos, (b[0], b[1], b[2], 1, 0), S, N)
        if merged is None and len(teams) >= 2:
            a, b = teams[-2], teams[-1]
            S, N = combined(a, b)
            form = max(a[1], b[1]) if a[1] != 3 and b[1] != 3 else 3
            if form == 1 and table_lds_bytes(1, S, N) + red > LDS_BYTES:
                form = 2
            if form != 3 and table_lds_bytes(form, S, N) + red > LDS_BYTES:
                form = 3 if has_pos else None
            if form is not None:
                merged = (teams, (b[0], form, 1, max(a[3], b[3]), min(a[4], b[4])), S, N)
        if merged is None:
            break
        kind, entry, S, N = merged
        kind[-2:] = [entry]
        sn[entry[0]] = (S, N)
    return teams + solos


def tune(trajs: np.ndarray, sizes: dict, has_pos: bool = True, sample: int = 1024,
         sample_hops: int = 2048, edges=BUCKET_EDGES, cache_path: str | None = CACHE_PATH,
         retune: bool = False, log=None) -> list:
    """Launch plan [(s_max, form, W, 0), ...] for the batch `trajs`.

    sizes: layout id -> (S, N).  Every bucket that holds trajectories is timed on
    up to `sample` of them with `sample_hops` hops each (or read from the cache).
    """
    device = engine.device_info()["name"]
    cache = _load_cache(cache_path) if cache_path else {}
    lay = trajs["layout"]
    S_of = np.array([sizes[int(x)][0] for x in lay])
    N_of = np.array([sizes[int(x)][1] for x in lay])
    bucket = np.array([bucket_of(int(s), edges) for s in S_of])
    plan = []
    changed = False
    for b, edge in enumerate(edges):
        idx = np.nonzero(bucket == b)[0]
        if idx.size == 0:
            continue
        tpl = idx.size / max(1, len(np.unique(lay[idx])))
        key = cache_key(device, edge, tpl)
        hit = cache.get(key)
        if hit and not retune:
            plan.append((edge, int(hit["form"]), int(hit["waves"]), 1))
            continue
        rng = np.random.RandomState(b)
        pick = idx if idx.size <= sample else np.sort(rng.choice(idx, sample, replace=False))
        t = trajs[pick].copy()
        t["prehops"] = 0
        t["hops"] = sample_hops
        smax, nmax = int(S_of[pick].max()), int(N_of[pick].max())
        # the bucket's own largest network fixes the LDS footprint of table forms
        smax_all, nmax_all = int(S_of[idx].max()), int(N_of[idx].max())
        scores = {}
        for form, w in candidates(max(smax, smax_all), max(nmax, nmax_all), has_pos):
            scores[(form, w)] = time_candidate(t, edge, form, w)
            if log:
                log(f"  bucket S<={edge}: form {form} W={w}: {scores[(form, w)]:.3e} hops/s")
        (form, w), rate = max(scores.items(), key=lambda kv: kv[1])
        cache[key] = {"form": form, "waves": w, "hops_per_s": rate,
                      "scores": {f"{f}/{ww}": v for (f, ww), v in scores.items()}}
        changed = True
        plan.append((edge, form, w, 1))
    if changed and cache_path:
        _save_cache(cache_path, cache)
    return plan
