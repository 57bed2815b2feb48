"""The CPU oracle against closed forms: Boltzmann statistics and the SET current.

These pin the restated hop loop independently of the reference's numbers: a
dropped Coulomb term, a wrong sign of dE or a transposed rate table would bias
the occupations away from exp(-H'/kT) or the current away from the exact
two-state value.
"""
import math

import numpy as np

import oracle
from tests import physics


def _prehop_occupation(net, hops, seed):
    """Time-weighted <n_i> with each waiting time given to the pre-hop state,
    rebuilt from the numba-loop oracle's event log and replayed uniforms."""
    N, P = net["N"], net["P"]
    U = oracle.stream_uniforms(seed, 0, hops).ravel()
    r = oracle.numba_simulate(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                              net["E_constant"], net["site_energies"],
                              net["transitions_constant"], net["occupation"],
                              np.zeros(P, np.int64), hops, uniforms=U, log=True)
    assert r.status == 0
    S = N + P
    occ = np.array(net["occupation"], dtype=float)
    acc = np.zeros(N)
    T = 0.0
    for h in range(hops):
        dt = -math.log(1.0 - U[2 * h]) * (1 / r.totals[h])
        acc += dt * occ
        T += dt
        f, t = divmod(int(r.events[h]), S)
        if f < N:
            occ[f] = 0
        if t < N:
            occ[t] = 1
    return acc / T


def test_grand_canonical_zero_bias_boltzmann():
    net = physics.boltzmann_network()
    exact = physics.boltzmann_site_occupation(net["distances"], net["E_constant"], net["I_0"],
                                              net["R"], V_e=net["V_e"])
    runs = np.array([_prehop_occupation(net, 60000, s) for s in range(6)])
    m = runs.mean(0)
    se = runs.std(0, ddof=1) / np.sqrt(len(runs))
    assert np.all(np.abs(m - exact) < 4 * se + 0.01), (m, exact, se)


def test_canonical_boltzmann_like_reference_validation():
    # validation/boltzmann_validation.py: N=10, M=0, no electrodes, n=5 carriers
    net = physics.boltzmann_network(seed=5, N=10, P=0, I_0=4.0, a=0.4, M=0)
    net["occupation"] = np.array([1] * 5 + [0] * 5, bool)
    exact = physics.boltzmann_site_occupation(net["distances"], net["E_constant"], net["I_0"],
                                              net["R"], n_fixed=5)
    runs = np.array([_prehop_occupation(net, 40000, s) for s in range(6)])
    m = runs.mean(0)
    se = runs.std(0, ddof=1) / np.sqrt(len(runs))
    assert abs(m.sum() - 5) < 1e-9
    assert np.all(np.abs(m - exact) < 4 * se + 0.01), (m, exact, se)


def test_set_current_exact():
    for E_G, V_S in [(0.0, 2.0), (1.0, -1.5), (-0.5, 0.7)]:
        net = physics.set_network(E_G, V_S)
        d = net["distances"]
        exact = physics.set_current(E_G, V_S, 0.0, net["transitions_constant"][1, 0],
                                    net["transitions_constant"][0, 2])
        eo, t, _, st = oracle.numba_ensemble(1, 2, 1.0, 1.0, net["I_0"], net["R"], d,
                                             net["E_constant"], net["site_energies"],
                                             net["transitions_constant"], net["occupation"],
                                             np.arange(16), 20000)
        assert (st == 0).all()
        cur = eo[:, 1] / t
        se = cur.std(ddof=1) / 4
        assert abs(cur.mean() - exact) < 4 * se + 1e-3 * abs(exact), (E_G, V_S, cur.mean(), exact)


def test_go_and_numba_restatements_agree():
    # the reference's own cross-engine check (thesis_indrek/generate_tests.py:109-144)
    net = physics.boltzmann_network(seed=9, N=12, P=4, V_e=0.0, I_0=5.0, a=0.3)
    net["site_energies"][12:] = [1.5, -1.5, 0.8, -0.8]
    args = (12, 4, 1.0, 1.0, net["I_0"], net["R"], net["distances"], net["E_constant"],
            net["site_energies"], net["transitions_constant"])
    eo1, t1, _, _ = oracle.numba_ensemble(*args, net["occupation"], np.arange(16), 20000)
    eo2, t2, _, _ = oracle.go_ensemble(*args, np.zeros(12, np.uint8), np.arange(16) + 100, 20000)
    c1, c2 = eo1 / t1[:, None], eo2 / t2[:, None]
    se = np.sqrt(c1.var(0, ddof=1) / 16 + c2.var(0, ddof=1) / 16)
    assert np.all(np.abs(c1.mean(0) - c2.mean(0)) < 4.5 * se + 1e-9)


def test_go_restatement_prunes_small_prefactors():
    """simulation.go:199-215: pairs with tc <= prune * max(tc) leave the pair list,
    so the logged (from, to) events never use one; with prune = 0 they do occur."""
    from tests.netgen import random_network, oracle_args
    net = random_network(21, 12, 4, I_0=50.0, a=0.4, vmax=1.0)
    tc = net["transitions_constant"]
    prune = 0.05
    pruned = tc <= prune * tc.max()
    np.fill_diagonal(pruned, False)
    assert 0.05 < pruned.mean() < 0.95
    args = oracle_args(net)
    hits = {}
    for p in (0.0, prune):
        r = oracle.go_simulate(*args, occupation=net["occupation"], hops=20_000,
                               prune_threshold=p, seed=3, log=True)
        assert r.status == 0
        ev = r.events
        hits[p] = int(pruned[ev[:, 0], ev[:, 1]].sum())
    assert hits[prune] == 0 and hits[0.0] > 0
