"""Error of the FP32 batched kernel's currents against the FP64 dense path.

python -m kmc_dn_amd.validate [--quick]

* ``ensemble``: the 4096-trajectory N=30 workload (1e5 hops per trajectory),
  both kernels on the same Philox streams; per-electrode mean current, standard
  error over trajectories, |z| and relative difference.
* ``fixtures``: the reference's own test networks (thesis_indrek fixtures, bias
  up to +-300 kT, large currents), 1e6 hops per trajectory: relative
  difference of every current whose magnitude exceeds 10 standard errors.
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

from . import _lib, engine, tables, workloads


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
    if path is None:
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "tests", "golden", "kmc_fixtures.npz")
    fx = np.load(path)
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
    rels, zs = [], []
    for i in range(n_layouts):
        m32, m64, se, z, rel = compare(out[_lib.FP32_BATCHED][i], out[_lib.FP64_DENSE][i])
        big = np.abs(m64) > 10 * se
        rels.append(rel[big])
        zs.append(z)
    rels, zs = np.concatenate(rels), np.concatenate(zs)
    return {"workload": f"{n_layouts} reference fixture networks x {replicas} trajectories x "
                        f"{hops} hops (+1e4 prehops), FP32 vs FP64 on the same streams",
            "currents_compared": int(rels.size), "max_rel": float(rels.max()),
            "median_rel": float(np.median(rels)), "frac_within_1pct": float(np.mean(rels < 0.01)),
            "max_z": float(zs.max()), "median_z": float(np.median(zs))}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args()
    engine.init(0)
    out = {"ensemble": ensemble_check(1024 if a.quick else 4096, 20000 if a.quick else 100_000),
           "fixtures": fixtures_check(4 if a.quick else 10, 8 if a.quick else 128,
                                      100_000 if a.quick else 1_000_000)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
