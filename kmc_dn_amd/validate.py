"""Error of the FP32 batched kernel's currents against the FP64 dense path, and
the reference's own accuracy statistic on its fixtures.

python -m kmc_dn_amd.validate [--quick] [--edb]

* ``ensemble``: the 4096-trajectory N=30 workload (1e5 hops per trajectory),
  both kernels on the same Philox streams; per-electrode mean current, standard
  error over trajectories, |z| and relative difference.
* ``fixtures``: the reference's own test networks (thesis_indrek fixtures, bias
  up to +-300 kT, large currents), 1e6 hops per trajectory: relative
  difference of every current whose magnitude exceeds 10 standard errors.
* ``edb``: the fixture harness of thesis_indrek/validate_tests.py:80-135 and
  :299-346 -- every fixture network run 5 times, the mean and population std
  of each electrode current compared with the fixture's stored 5-run mean and
  std by the Bhattacharyya distance, averaged over fixtures and electrodes
  (E_DB) -- for this engine's back ends, beside the reference's published
  figures (thesis_indrek/printTables.py:116-136).
"""
from __future__ import annotations

import argparse
import json
import math
import os

import numpy as np

from . import _lib, engine, tables, workloads

FIXTURES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "tests", "golden", "kmc_fixtures.npz")

# thesis_indrek/printTables.py:114-136: E_DB of the reference's back ends against its
# fixture sets, two campaigns ("RND" = rnd_min_max networks, "XOR" = XOR_wide networks)
REFERENCE_EDB = {
    "numba 1e6": {"RND": (0.14, 0.30), "XOR": (0.34, 0.49)},
    "Go 1e6": {"RND": (0.55, 0.46), "XOR": (0.59, 2.5)},
    "Go recording 1e6": {"RND": (0.19, 0.34), "XOR": (0.36, 3.7)},
    "Go recording 5e6": {"RND": (0.30, 0.17), "XOR": (0.47, 0.20)},
    "Go pruning 1e-5": {"RND": (120, 340), "XOR": (58, 290)},
    "Go pruning 1e-7": {"RND": (0.91, 2.0), "XOR": (0.63, 2.2)},
    "Go pruning 1e-9": {"RND": (0.54, 0.40), "XOR": (0.68, 3.1)},
}


def mean_std_pop(runs):
    """Per-column mean and population std (divide by n) of a (runs, P) array,
    as getMeanAndStandardDeviation (validate_tests.py:17-28) computes them."""
    runs = np.asarray(runs, float)
    m = runs.mean(0)
    return m, np.sqrt(((runs - m) ** 2).mean(0))


def calc_d(mean_p, mean_v, std_p, std_v):
    """Bhattacharyya distance of two normal estimates (validate_tests.py:80-87);
    a zero variance on either side scores log(10000)."""
    var_p, var_v = std_p ** 2, std_v ** 2
    if var_p == 0 or var_v == 0:
        return math.log(10000)
    part1 = 0.25 * math.log(0.25 * (var_p / var_v + var_v / var_p + 2))
    part2 = 0.25 * (((mean_p - mean_v) ** 2) / (var_p + var_v))
    return part1 + part2


def e_db(ref_mean, ref_std, new_mean, new_std):
    """E_DB and sigma_DB (evaluate_BC, validate_tests.py:90-105): the mean and
    population std of calc_d over every (fixture, electrode); also the values."""
    d = [calc_d(a, b, c, e) for a, b, c, e in zip(np.ravel(ref_mean), np.ravel(new_mean),
                                                  np.ravel(ref_std), np.ravel(new_std))]
    d = np.asarray(d)
    return float(d.mean()), float(np.sqrt(((d - d.mean()) ** 2).mean())), d


def _run(layout, es, seeds, hops, occ, precision, prehops=0):
    b = engine.Batch(engine.make_trajs(len(seeds), layout.id, es, seeds, hops, occ,
                                       prehops=prehops))
    b.run(precision=precision)
    ms = b.sync()
    r = b.results()
    b.close()
    assert (r.status == 0).all()
    return r, ms


def compare(c32, c64):
    m32, m64 = c32.mean(0), c64.mean(0)
    se = np.sqrt(c32.var(0, ddof=1) / len(c32) + c64.var(0, ddof=1) / len(c64))
    z = np.abs(m32 - m64) / np.where(se > 0, se, np.inf)
    rel = np.abs(m32 - m64) / np.maximum(np.abs(m64), 1e-300)
    return m32, m64, se, z, rel


def ensemble_check(K=4096, hops=100_000):
    ens = workloads.ensemble(0, K, hops)
    n = ens.net
    lay = engine.Layout(n.N, n.P, n.nu, n.kT, n.I_0, n.R, n.distances, n.transitions_constant)
    es = lay.energy_sets(n.E_constant, n.site_energies[n.N:])[0]
    r32, ms32 = _run(lay, es, ens.seeds, hops, n.occupation, _lib.FP32_BATCHED)
    r64, ms64 = _run(lay, es, ens.seeds, hops, n.occupation, _lib.FP64_DENSE)
    m32, m64, se, z, rel = compare(r32.currents(n.P), r64.currents(n.P))
    return {"workload": f"{K} trajectories x {hops} hops, N=30 P=8 a=0.1 I_0=50",
            "fp32": m32.tolist(), "fp64": m64.tolist(), "se": se.tolist(), "z": z.tolist(),
            "rel": rel.tolist(), "fp32_hops_per_s": K * hops / ms32 * 1e3,
            "fp64_hops_per_s": K * hops / ms64 * 1e3,
            "electrode_events_per_traj": float(np.abs(r64.eo[:, :n.P]).sum(1).mean())}


def fixtures_check(n_layouts=10, replicas=128, hops=1_000_000, path=None):
    fx = np.load(path or FIXTURES)
    N, P = int(fx["N"]), int(fx["P"])
    trajs = []
    for i, k in enumerate(np.linspace(0, 99, n_layouts).astype(int)):
        lset = "rnd" if i % 2 == 0 else "xor"
        acc, el = fx[f"{lset}_acceptors"][k], fx[f"{lset}_electrodes"][k]
        d = tables.calc_distances(acc, el)
        tc = tables.calc_transitions_constant(d, float(fx["ab"]), 1.0)
        lay = engine.Layout(N, P, 1.0, 1.0, float(fx["I_0"]), float(fx["R"]), d, tc)
        es = lay.energy_sets(fx[f"{lset}_E_constant"][k], el[:, 3])[0]
        seeds = np.arange(replicas, dtype=np.uint64) + 1000 * int(k) + 100000 * i
        trajs.append(engine.make_trajs(replicas, lay.id, es, seeds, hops,
                                       fx[f"{lset}_occupation"][k], prehops=10_000))
    trajs = np.concatenate(trajs)
    out = {}
    for prec in (_lib.FP32_BATCHED, _lib.FP64_DENSE):
        b = engine.Batch(trajs)
        b.run(precision=prec)
        b.sync()
        r = b.results()
        b.close()
        assert (r.status == 0).all()
        out[prec] = r.currents(P).reshape(n_layouts, replicas, P)
    rels, zs, tight = [], [], []
    for i in range(n_layouts):
        m32, m64, se, z, rel = compare(out[_lib.FP32_BATCHED][i], out[_lib.FP64_DENSE][i])
        big = np.abs(m64) > 10 * se
        rels.append(rel[big])
        # a 1 % bar is testable at 3 sigma where 3 standard errors are below 1 %
        tight.append(rel[np.abs(m64) > 300 * se])
        zs.append(z)
    rels, zs, tight = np.concatenate(rels), np.concatenate(zs), np.concatenate(tight)
    return {"workload": f"{n_layouts} reference fixture networks x {replicas} trajectories x "
                        f"{hops} hops (+1e4 prehops), FP32 vs FP64 on the same streams",
            "currents_compared": int(rels.size), "max_rel": float(rels.max()),
            "median_rel": float(np.median(rels)), "frac_within_1pct": float(np.mean(rels < 0.01)),
            "currents_1pct_testable": int(tight.size),
            "max_rel_1pct_testable": float(tight.max()) if tight.size else None,
            "max_z": float(zs.max()), "median_z": float(np.median(zs))}


# fixture set -> the layout family it was measured on (tests/golden/kmc_fixtures.npz)
FIXTURE_SETS = {"rnd_min_max": "rnd", "XOR_wide": "xor", "rnd_min_max5M": "rnd",
                "XOR_wide5M": "xor"}


def edb_runs(fixture_sets, hops, runs=5, n=100, precision=_lib.FP32_BATCHED, prune=0.0,
             start="empty", seed0=0, path=None):
    """Per-run currents (len(fixture_sets), n, runs, P) of the fixture networks.

    start="empty": each run starts from an empty network, as go_simulation's
    wrappers do (simulationWrapper.go:134-141); start="chain": the runs of a
    network continue one another from the fixture's stored occupation, as
    repeated python_simulation calls do (kmc_dopant_networks.py:513-542).  All
    networks and runs of one call go into one batch (runs of a chain as
    consecutive hop ranges of one trajectory, read after each range)."""
    fx = np.load(path or FIXTURES)
    N, P = int(fx["N"]), int(fx["P"])
    layouts, trajs = [], []
    for si, name in enumerate(fixture_sets):
        fam = FIXTURE_SETS[name]
        for k in range(n):
            acc, el = fx[f"{fam}_acceptors"][k], fx[f"{fam}_electrodes"][k]
            d = tables.calc_distances(acc, el)
            tc = tables.calc_transitions_constant(d, float(fx["ab"]), float(fx["nu"]))
            lay = engine.Layout(N, P, float(fx["nu"]), float(fx["kT"]),
                                float(fx["I_0"]) * float(fx["kT"]), float(fx["R"]), d, tc, prune)
            layouts.append(lay)
            es = lay.energy_sets(fx[f"{fam}_E_constant"][k], el[:, 3])[0]
            base = seed0 + 1_000_000 * si + 1000 * k
            if start == "empty":
                trajs.append(engine.make_trajs(runs, lay.id, es,
                                               np.arange(base, base + runs, dtype=np.uint64),
                                               hops, np.zeros(N, bool)))
            else:
                trajs.append(engine.make_trajs(1, lay.id, es, base, runs * hops,
                                               fx[f"{fam}_occupation"][k]))
    trajs = np.concatenate(trajs)
    out = []
    try:
        b = engine.Batch(trajs)
        if start == "empty":
            b.run(precision=precision)
            b.sync()
            r = b.results()
            assert (r.status == 0).all()
            cur = r.currents(P)
        else:   # consecutive ranges of one trajectory per network (FP32 resume)
            if precision != _lib.FP32_BATCHED:
                raise ValueError("chained runs use the FP32 kernel's resume ranges")
            cur = np.zeros((len(trajs), runs, P))
            for j in range(runs):
                b.set_state(h_begin=np.full(len(trajs), j * hops),
                            h_end=np.full(len(trajs), (j + 1) * hops),
                            time=np.zeros(len(trajs)) if j else None,
                            eo=np.zeros((len(trajs), _lib.PMAX), np.int64) if j else None)
                b.run(precision=precision, resume=True)
                b.sync()
                r = b.results()
                assert (r.status == 0).all()
                cur[:, j] = r.currents(P)
        b.close()
    finally:
        for lay in layouts:
            lay.close()
    return cur.reshape(len(fixture_sets), n, runs, P)


def edb_table(fixture_sets, cur, path=None):
    """E_DB, sigma_DB of per-run currents (from edb_runs) against each set."""
    fx = np.load(path or FIXTURES)
    out = {}
    for si, name in enumerate(fixture_sets):
        n = cur.shape[1]
        m = np.zeros((n, cur.shape[3]))
        sd = np.zeros_like(m)
        for k in range(n):
            m[k], sd[k] = mean_std_pop(cur[si, k])
        e, s, _ = e_db(fx[f"{name}_mean_currents"][:n], fx[f"{name}_stddev_currents"][:n], m, sd)
        out[name] = (e, s)
    return out


def edb_check(quick=False):
    """The reference's fixture harness for this engine's back ends."""
    n = 20 if quick else 100
    sets_1m, sets_5m = ["rnd_min_max", "XOR_wide"], ["rnd_min_max5M", "XOR_wide5M"]
    rows = {}

    def row(label, sets, **kw):
        cur = edb_runs(sets, n=n, **kw)
        rows[label] = {k: [round(v[0], 3), round(v[1], 3)] for k, v in edb_table(sets, cur).items()}

    row("fp32 go_simulation 1e6 (empty start)", sets_1m + sets_5m, hops=1_000_000)
    row("fp32 go_simulation 5e6 (empty start)", sets_1m + sets_5m, hops=5_000_000)
    for pr in (1e-5, 1e-7, 1e-9):
        row(f"fp32 pruned {pr:g} 1e6", sets_1m, hops=1_000_000, prune=pr)
    row("fp32 chained runs 1e6 (python_simulation start)", sets_1m, hops=1_000_000,
        start="chain")
    return {"networks_per_set": n, "runs": 5, "rows": rows,
            "reference_printTables": REFERENCE_EDB}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--edb", action="store_true", help="only the fixture E_DB harness")
    a = ap.parse_args()
    engine.init(0)
    if a.edb:
        print(json.dumps({"edb": edb_check(a.quick)}))
        return
    out = {"ensemble": ensemble_check(1024 if a.quick else 4096, 20000 if a.quick else 100_000),
           "fixtures": fixtures_check(4 if a.quick else 10, 8 if a.quick else 128,
                                      100_000 if a.quick else 1_000_000),
           "edb": edb_check(a.quick)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
