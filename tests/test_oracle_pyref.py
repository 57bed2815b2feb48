"""The C oracle against the pure-Python restatement, bit for bit, on small networks.

Both follow kmc_dopant_networks.py:33-163 independently; with the same replayed
uniforms every event index, the clock, the tallies and the record arrays must
be identical.
"""
import numpy as np
import pytest

import oracle
from oracle import pyref
from tests.netgen import random_network


@pytest.mark.parametrize("seed,N,P", [(1, 3, 2), (2, 5, 0), (3, 6, 3), (4, 8, 4), (5, 2, 1),
                                      (6, 7, 8)])
def test_c_oracle_matches_python(seed, N, P):
    net = random_network(seed, N, P, M=1 if N > 2 else 0, I_0=20.0, a=0.4, vmax=3.0)
    hops = 250
    rng = np.random.RandomState(100 + seed)
    U = rng.rand(2 * hops)
    U[7] = 0.0          # an event uniform of exactly 0 selects pair (0, 0)
    occ = net["occupation"]
    if P == 0:
        occ = np.zeros(N, bool)
        occ[: N // 2] = True
    eo = np.zeros(P, np.int64)
    try:
        py = pyref.simulate(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"].tolist(),
                            net["E_constant"].tolist(), net["site_energies"].tolist(),
                            net["transitions_constant"].tolist(), occ.tolist(), eo.tolist(), hops,
                            U.tolist(), record=True)
    except ZeroDivisionError:
        pytest.skip("network froze (total rate 0)")
    r = oracle.numba_simulate(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                              net["E_constant"], net["site_energies"], net["transitions_constant"],
                              occ, eo, hops, record=True, uniforms=U, log=True)
    t, o, e, tr, oit, ev = py
    assert r.status == 0
    assert list(r.events) == ev
    assert r.events[3] == 0          # U[7] is the event uniform of hop 3
    assert r.time == t
    assert list(r.electrode_occupation) == e
    assert list(r.occupation) == o
    assert np.array_equal(r.traffic, np.array(tr))
    assert np.array_equal(r.occupations_in_time, np.array(oit))


def test_zero_total_rate_is_an_error():
    # no electrodes and every acceptor occupied: no possible hop (reference: ZeroDivisionError)
    net = random_network(9, 5, 0, M=0)
    occ = np.ones(5, bool)
    with pytest.raises(ZeroDivisionError):
        pyref.simulate(5, 0, 1.0, 1.0, net["I_0"], net["R"], net["distances"].tolist(),
                       net["E_constant"].tolist(), net["site_energies"].tolist(),
                       net["transitions_constant"].tolist(), occ.tolist(), [], 3, [0.5] * 6)
    r = oracle.numba_simulate(5, 0, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                              net["E_constant"], net["site_energies"], net["transitions_constant"],
                              occ, np.zeros(0, np.int64), 3, uniforms=np.full(6, 0.5))
    assert r.status == 2


def test_philox_stream_equals_replayed_stream():
    net = random_network(11, 10, 4)
    U = oracle.stream_uniforms(42, 17, 500).ravel()
    args = (10, 4, 1.0, 1.0, net["I_0"], net["R"], net["distances"], net["E_constant"],
            net["site_energies"], net["transitions_constant"], net["occupation"],
            np.zeros(4, np.int64), 500)
    a = oracle.numba_simulate(*args, uniforms=U, log=True)
    b = oracle.numba_simulate(*args, seed=42, hop0=17, log=True)
    assert np.array_equal(a.events, b.events) and a.time == b.time
