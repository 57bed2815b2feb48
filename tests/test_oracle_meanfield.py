"""Mean-field estimators of the Go library, restated in the oracle:
probSimulate (goSimulation/probabilitySimulation.go:53-157) and
calcAverageCurrent (goSimulation/simulation.go:671-736).

Pins: the C restatement equals an independent pure-Python reading bit for bit;
the integrator conserves holes; on one acceptor between two electrodes both
reduce to the exact single-electron-transistor current (the mean field of one
site is its exact master equation), up to the reference's 1/0.98 factor between
moved occupation and advanced time."""
import numpy as np
import pytest

import oracle
from oracle import pyref
from tests import physics
from tests.netgen import random_network


def _args(net):
    return (net["N"], net["P"], net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
            net["E_constant"], net["site_energies"], net["transitions_constant"])


@pytest.mark.parametrize("seed,N,P", [(3, 6, 3), (4, 8, 0), (5, 5, 4), (6, 3, 2)])
def test_prob_simulate_c_equals_python(seed, N, P):
    net = random_network(seed, N, P, M=2, I_0=20.0, a=0.3, vmax=1.0)
    occ0 = np.full(N, 0.5)
    r = oracle.prob_simulate(*_args(net), occ0, 40, record=True)
    py = pyref.prob_simulate(N, P, net["nu"], net["kT"], net["I_0"], net["R"], list(occ0),
                             net["distances"].tolist(), list(net["E_constant"]),
                             net["transitions_constant"].tolist(), list(net["site_energies"]), 40,
                             record=True)
    assert r.time == py[0]
    assert np.array_equal(r.occupation, py[1])
    assert np.array_equal(r.electrode_occupation, py[2])
    assert np.array_equal(r.site_energies[:N], py[3][:N])
    assert np.array_equal(r.traffic, np.array(py[4]))
    assert np.array_equal(r.average_occupation, py[5])


def test_prob_simulate_conserves_holes():
    """The step is capped so that no occupation leaves [0, 1]: what the acceptors
    lose the electrodes gain, and traffic is antisymmetric."""
    net = random_network(7, 20, 6, I_0=50.0, a=0.2, vmax=2.0)
    occ0 = np.full(20, 0.5)
    r = oracle.prob_simulate(*_args(net), occ0, 500, record=True)
    assert ((r.occupation >= 0) & (r.occupation <= 1)).all()
    assert abs(r.occupation.sum() + r.electrode_occupation.sum() - occ0.sum()) < 1e-10
    assert np.allclose(r.traffic, -r.traffic.T, atol=1e-12)
    assert np.allclose(r.traffic[:, 20:].sum(0), r.electrode_occupation, atol=1e-10)
    assert r.time > 0 and (r.average_occupation > 0).all()


@pytest.mark.parametrize("E_G,V_S", [(0.3, 1.0), (-0.4, 2.0), (1.5, 0.5)])
def test_single_site_reduces_to_the_exact_set_current(E_G, V_S):
    net = physics.set_network(E_G, V_S, ab_over_R=2.0)
    exact = physics.set_current(E_G, V_S, 0.0, net["transitions_constant"][1, 0],
                                net["transitions_constant"][0, 2])
    # a step moves the occupation by the net flow over the total rate, a full
    # unit: the iterate settles into a two-step cycle between near 0 and near 1
    # whose currents, over an even number of steps, are the stationary ones
    warm = oracle.prob_simulate(*_args(net), np.array([0.5]), 1000)
    r = oracle.prob_simulate(*_args(net), warm.occupation, 1000)
    # occupation moves by rate * step while the clock advances 0.98 * step
    assert r.electrode_occupation[1] / r.time * 0.98 == pytest.approx(exact, rel=1e-9)
    assert r.electrode_occupation[0] / r.time * 0.98 == pytest.approx(-exact, rel=1e-9)
    # calcAverageCurrent at the stationary occupation of the master equation,
    # p = (G_SA + G_DA) / (G_SA + G_DA + G_AS + G_AD): the same current / 0.98
    tcS, tcD = net["transitions_constant"][1, 0], net["transitions_constant"][0, 2]
    G_SA, G_AS = tcS * min(1.0, np.exp(-(E_G - V_S))), tcS * min(1.0, np.exp(-(V_S - E_G)))
    G_AD, G_DA = tcD * min(1.0, np.exp(E_G)), tcD * min(1.0, np.exp(-E_G))
    p = (G_SA + G_DA) / (G_SA + G_DA + G_AS + G_AD)
    cur = oracle.average_current(*_args(net), np.array([p]))
    assert cur[1] * 0.98 == pytest.approx(exact, rel=2e-6)
    assert cur[0] * 0.98 == pytest.approx(-exact, rel=2e-6)
