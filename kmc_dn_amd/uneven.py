"""Uneven batches: thousands of distinct networks in one launch.

Each network of an :class:`~kmc_dn_amd.workloads.UnevenBatch` becomes a device
layout (tables plus site positions) with one energy set; all trajectories go
into one batch whose size buckets run under the autotuned launch plan, the
bucket kernels forked from one stream and captured in one hipGraph.

The per-network physics is the reference's (kmc_dopant_networks.py:33-135);
batching many networks per call has no reference counterpart (the reference
runs one network per call, simulationWrapper.go:83-169).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import autotune, engine


@dataclass
class UnevenDevice:
    layouts: list            # engine.Layout per network
    esets: np.ndarray        # energy-set id per network
    trajs: np.ndarray        # TRAJ_DTYPE per network
    sizes: dict              # layout id -> (S, N)

    def close(self) -> None:
        for lay in self.layouts:
            lay.close()
        self.layouts = []


def upload(ub, prehops: int = 0) -> UnevenDevice:
    """Layouts, positions and energy sets of every network; one trajectory each
    (Philox key = its seed, initial occupation = the network's)."""
    lays, esets, sizes = [], [], {}
    for net in ub.nets:
        lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                            net.transitions_constant)
        pos = np.concatenate([net.acceptors, net.electrodes[:, :3]], 0)
        lay.set_positions(net.ab, pos)
        esets.append(int(lay.energy_sets(net.E_constant, net.site_energies[net.N:])[0]))
        lays.append(lay)
        sizes[lay.id] = (lay.S, lay.N)
    n = len(ub.nets)
    t = np.zeros(n, dtype=engine.TRAJ_DTYPE)
    t["layout"] = [lay.id for lay in lays]
    t["eset"] = esets
    t["seed"] = ub.seeds
    t["prehops"] = prehops
    t["hops"] = ub.hops
    t["occ"] = np.stack([engine.occupation_to_bits(net.occupation) for net in ub.nets])
    return UnevenDevice(lays, np.array(esets, dtype=np.int32), t, sizes)


@dataclass
class UnevenRun:
    result: engine.BatchResult
    plan: list
    kernel_ms: float
    wall_s: float
    hops: int

    @property
    def hops_per_s(self) -> float:
        return self.hops / (self.kernel_ms * 1e-3)


def run(dev: UnevenDevice, plan=None, use_graph: bool = True, tune_log=None,
        recompute_interval: int = 1024, repeats: int = 1) -> UnevenRun:
    """Run every trajectory of `dev` to its budget under `plan` (autotuned when
    None); the kernel time is the HIP-event time of the last of `repeats` runs."""
    if plan is None:
        plan = autotune.plan_for(dev.trajs, dev.sizes, log=tune_log)
    b = engine.Batch(dev.trajs)
    try:
        b.set_buckets(plan)
        ms = 0.0
        t0 = time.perf_counter()
        for _ in range(repeats):
            t0 = time.perf_counter()
            b.run(precision=L.FP32_BATCHED, use_graph=use_graph,
                  recompute_interval=recompute_interval)
            ms = b.sync()
        wall = time.perf_counter() - t0
        res = b.results()
    finally:
        b.close()
    hops = int(np.sum(dev.trajs["hops"] + dev.trajs["prehops"]))
    return UnevenRun(res, list(plan), ms, wall, hops)


def check_invariants(dev: UnevenDevice, res: engine.BatchResult) -> None:
    """Status, budgets, time, and charge conservation of every trajectory:
    sum_e eo_e = (holes filled) - (holes emptied) = occupied_end - occupied_start."""
    if (res.status != 0).any():
        bad = np.nonzero(res.status != 0)[0]
        raise AssertionError(f"{bad.size} trajectories stopped early (first {bad[:5]})")
    if not (res.hops_done == dev.trajs["hops"]).all():
        raise AssertionError("hop budgets not met")
    if not (res.time > 0).all():
        raise AssertionError("non-positive clock")
    n0 = np.array([int(sum(bin(int(w)).count("1") for w in row)) for row in dev.trajs["occ"]])
    n1 = np.array([int(sum(bin(int(w)).count("1") for w in row)) for row in res.occ_bits])
    if not (res.eo.sum(1) == -(n1 - n0)).all():
        raise AssertionError("electrode tallies do not balance the occupation change")
