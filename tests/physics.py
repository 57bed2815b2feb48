"""Closed forms the hop loop must reproduce (exact enumeration, two-state chain)."""
from __future__ import annotations

import itertools

import numpy as np

from kmc_dn_amd import tables as T


def boltzmann_site_occupation(distances, E_constant, I_0, R, V_e=None, n_fixed=None, kT=1.0):
    """<n_i> under p(n) ~ exp(-H'(n)/kT) by enumeration over all microstates.

    H'(n) = I_0 R sum_{i<j} e_i e_j / d_ij - sum_i e_i E_i - V_e sum_i n_i, e = 1 - n.
    This is the energy whose differences are the hop loop's dE
    (kmc_dopant_networks.py:59-84): for an acceptor pair dE = se_j - se_i - I_0R/d_ij
    and for electrode e -> acceptor j dE = se_j - V_e, so Miller-Abrahams rates
    obey detailed balance with respect to it.  With n_fixed (no electrodes) the
    ensemble is canonical; V_e is then irrelevant (kmc_dopant_networks.py:982-1001
    is this H without the donor term, which is constant for fixed n).
    """
    N = len(E_constant)
    d = np.asarray(distances)[:N, :N]
    inv = np.zeros((N, N))
    iu = np.triu_indices(N, 1)
    inv[iu] = 1.0 / d[iu]
    states = np.array(list(itertools.product([0, 1], repeat=N)), dtype=float)
    if n_fixed is not None:
        states = states[states.sum(1) == n_fixed]
    e = 1.0 - states
    H = I_0 * R * np.einsum("si,ij,sj->s", e, inv, e) - e @ np.asarray(E_constant)
    if V_e is not None and n_fixed is None:
        H = H - V_e * states.sum(1)
    w = np.exp(-(H - H.min()) / kT)
    p = w / w.sum()
    return p @ states


def set_current(E_G, V_S, V_D, tc_SG, tc_GD, nu=1.0, kT=1.0):
    """Exact stationary drain current of one acceptor between two electrodes.

    Rates Gamma_xy = nu * tc_xy * min(1, exp(-(E_y - E_x)/kT)) (kmc_dopant_networks.py:69-87;
    the table tc already carries one nu).  Holes arriving at the drain count +1:
    I_D = (G_SG G_GD - G_GS G_DG) / (G_SG + G_GS + G_GD + G_DG).
    """
    def g(tc, Ex, Ey):
        return nu * tc * min(1.0, np.exp(-(Ey - Ex) / kT))
    SG, GS = g(tc_SG, V_S, E_G), g(tc_SG, E_G, V_S)
    GD, DG = g(tc_GD, E_G, V_D), g(tc_GD, V_D, E_G)
    return (SG * GD - GS * DG) / (SG + GS + GD + DG)


def boltzmann_network(seed=3, N=8, P=2, V_e=0.4, I_0=3.0, a=0.35, M=1):
    rng = np.random.RandomState(seed)
    acc = rng.rand(N, 3) * [1, 1, 0]
    don = rng.rand(M, 3) * [1, 1, 0]
    el = np.zeros((P, 4))
    for e in range(P):
        el[e] = [(e + 0.5) / P, 0.0 if e % 2 == 0 else 1.0, 0.0, V_e]
    _, R = T.characteristic_length(N, 1.0, 1.0, 0.0)
    d = T.calc_distances(acc, el)
    tc = T.calc_transitions_constant(d, a * R, 1.0)
    E = rng.uniform(-1.5, 1.5, N) + T.calc_comp_constant(acc, don, I_0, R)
    se = np.zeros(N + P)
    se[N:] = V_e
    occ = np.zeros(N, bool)
    occ[: N // 2] = True
    return dict(N=N, P=P, nu=1.0, kT=1.0, I_0=I_0, R=R, distances=d, transitions_constant=tc,
                E_constant=E, site_energies=se, occupation=occ, V_e=V_e)


def set_network(E_G, V_S, V_D=0.0, ab_over_R=1e5):
    """Single-acceptor transistor of validation/set/set.py:33-59 (1D, acceptor at x/2)."""
    acc = np.array([[0.5, 0.0, 0.0]])
    el = np.array([[0.0, 0.0, 0.0, V_S], [1.0, 0.0, 0.0, V_D]])
    R = 1.0   # N / xdim = 1
    d = T.calc_distances(acc, el)
    tc = T.calc_transitions_constant(d, ab_over_R * R, 1.0)
    se = np.array([0.0, V_S, V_D])
    return dict(N=1, P=2, nu=1.0, kT=1.0, I_0=100.0, R=R, distances=d, transitions_constant=tc,
                E_constant=np.array([E_G]), site_energies=se, occupation=np.array([False]))
