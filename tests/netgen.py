"""Test networks: the reference's fixture layouts and seeded synthetic ones."""
from __future__ import annotations

import numpy as np

from kmc_dn_amd import tables as T


def fixture_network(fx: dict, layout_set: str, k: int) -> dict:
    """Network k of the reference's 'rnd' or 'xor' fixture layouts (N=30, M=3, P=8)."""
    N, P = int(fx["N"]), int(fx["P"])
    acc = fx[f"{layout_set}_acceptors"][k]
    el = fx[f"{layout_set}_electrodes"][k]
    d = T.calc_distances(acc, el)
    tc = T.calc_transitions_constant(d, float(fx["ab"]), float(fx["nu"]))
    se = np.zeros(N + P)
    se[N:] = el[:, 3]
    return dict(N=N, P=P, nu=float(fx["nu"]), kT=float(fx["kT"]),
                I_0=float(fx["I_0"]) * float(fx["kT"]), R=float(fx["R"]), distances=d,
                transitions_constant=tc, E_constant=fx[f"{layout_set}_E_constant"][k].copy(),
                site_energies=se, occupation=fx[f"{layout_set}_occupation"][k].copy())


def random_network(seed: int, N: int, P: int, M: int | None = None, I_0: float = 50.0,
                   a: float = 0.1, vmax: float = 1.0, dim: int = 2, zdim: float = 0.0,
                   E_spread: float = 1.0) -> dict:
    """Seeded network in the unit square (or a 1x1xzdim box), perimeter electrodes.

    E_constant = random site energies (no Laplace solve) + donor compensation.
    """
    rng = np.random.RandomState(seed)
    M = max(0, round(N / 10)) if M is None else M
    z = zdim if dim == 3 else 0.0
    acc, don = T.place_dopants_random(rng, N, M, 1.0, 1.0, z)
    el = np.zeros((P, 4))
    for e in range(P):   # evenly spaced on the perimeter
        s = (e + 0.5) / P * 4.0
        side, f = int(s), s - int(s)
        el[e, :3] = [(f, 0, z / 2), (1, f, z / 2), (1 - f, 1, z / 2), (0, 1 - f, z / 2)][side]
    el[:, 3] = rng.uniform(-vmax, vmax, P)
    _, R = T.characteristic_length(N, 1.0, 1.0, z)
    ab = a * R
    d = T.calc_distances(acc, el)
    tc = T.calc_transitions_constant(d, ab, 1.0)
    comp = T.calc_comp_constant(acc, don, I_0, R)
    E = rng.uniform(-E_spread, E_spread, N) + comp
    se = np.zeros(N + P)
    se[N:] = el[:, 3]
    occ = T.place_charges_random(rng, N, M)
    return dict(N=N, P=P, nu=1.0, kT=1.0, I_0=I_0, R=R, distances=d, transitions_constant=tc,
                E_constant=E, site_energies=se, occupation=occ, acceptors=acc, electrodes=el)


def oracle_args(net: dict) -> tuple:
    return (net["N"], net["P"], net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
            net["E_constant"], net["site_energies"], net["transitions_constant"])
