"""IV curves: every voltage point and replica of a sweep in one device launch.

Mirrors ``IV(kmc_dn, electrode, voltagelist, tol, interval, hops, prehops)``
(kmc_dopant_networks_utils.py:639-693), which loops over the voltages, updates
the potential (linear scaling of eV_constant when the other electrodes are at
zero, :655-676, else a full Laplace solve) and calls go_simulation per point.
Here the potential of every point comes from the unit-voltage Laplace basis
(superposition, valid for any electrode voltages), each point becomes an
energy set of one shared layout, and all points x replicas run as independent
trajectories of a single batch.  Returns ``currentlist[P, n_voltages]`` like
the reference, plus the standard error over replicas.
"""
from __future__ import annotations

import numpy as np

from . import _lib, engine, laplace


def iv_batch(layout: engine.Layout, E_sets: np.ndarray, V_sets: np.ndarray, replicas: int,
             hops: int, occupation, prehops: int = 0, seed0: int = 0):
    """Upload the energy sets and build the batch (voltage-major, replicas inner).

    Trajectory k = i * replicas + r has Philox key seed0 + k.  Returns the batch and
    the energy-set ids (free them with engine.Layout.drop_energy_sets when done).
    """
    ids = layout.energy_sets(E_sets, V_sets)
    nV = len(ids)
    K = nV * replicas
    t = engine.make_trajs(K, layout.id, np.repeat(ids, replicas),
                          np.arange(seed0, seed0 + K, dtype=np.uint64), hops, occupation,
                          prehops=prehops)
    return engine.Batch(t), ids


def currents_from(res: engine.BatchResult, nV: int, replicas: int, P: int):
    cur = res.currents(P).reshape(nV, replicas, P)
    mean = cur.mean(axis=1).T
    se = (cur.std(axis=1, ddof=1) / np.sqrt(replicas)).T if replicas > 1 else np.zeros_like(mean)
    return mean, se


def IV(kmc, electrode: int, voltagelist, hops: int = 0, prehops: int = 0, replicas: int = 1,
       precision: int = _lib.FP32_BATCHED, seed0: int | None = None, return_stderr=False,
       start: str = "empty", **_):
    """currentlist[P, n] of a kmc_dn model with `electrode` swept over `voltagelist`.

    start="empty" (default) is the reference's behaviour: each point runs
    go_simulation(goSpecificFunction='wrapperSimulateRecord') (utils :683-684),
    whose Go wrapper starts from an empty network and tallies every hop
    (simulationWrapper.go:134-141); the reference's numba prehops move only the
    model's own occupation, which that run ignores, so `prehops` adds nothing
    here.  start="model" instead starts every point from the model's current
    occupation and runs `prehops` untallied hops before the `hops` tallied ones.
    """
    if start not in ("empty", "model"):
        raise ValueError("start must be 'empty' or 'model'")
    if hops <= 0:
        raise _lib.KmcError(_lib.ERR_UNSUPPORTED, "convergence-based IV (hops=0) is not supported")
    volts = np.asarray(voltagelist, dtype=np.float64)
    basis = laplace.LaplaceBasis(kmc.acceptors, kmc.electrodes, kmc.xdim, kmc.ydim, res=kmc.res,
                                 static_electrodes=kmc.static_electrodes)
    Vs = np.repeat(kmc.electrodes[None, :, 3], len(volts), axis=0)
    Vs[:, electrode] = volts
    E = basis.eV(Vs, kmc.mu) + kmc.comp_constant[None, :]
    lay = kmc._layout()
    seed = int(np.random.randint(0, 2 ** 31)) if seed0 is None else seed0
    if start == "empty":
        occ0, pre = np.zeros(kmc.N, dtype=bool), 0
    else:
        occ0, pre = kmc.occupation, int(prehops)
    b, ids = iv_batch(lay, E, Vs, replicas, int(hops), occ0, pre, seed)
    b.run(precision=precision)
    b.sync()
    r = b.results()
    b.close()
    engine.Layout.drop_energy_sets(ids)
    for st in np.unique(r.status):
        _lib.check(int(st))
    mean, se = currents_from(r, len(volts), replicas, kmc.P)
    kmc.electrodes[electrode, 3] = volts[-1]   # the reference leaves the last point applied
    kmc.update_V()
    return (mean, se) if return_stderr else mean
