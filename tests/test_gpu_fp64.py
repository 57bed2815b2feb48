"""FP64 dense kernel against the oracle (kmc_dopant_networks.py:33-135).

With a replayed uniform stream the event index sequence must equal the CPU
restatement's for the first 10^4 hops (bit-exact); with the Philox stream a
batch of trajectories must reproduce the oracle's ensemble trajectory by
trajectory.
"""
import numpy as np
import pytest

import oracle
from kmc_dn_amd import _lib, engine
from tests.netgen import fixture_network, random_network

pytestmark = pytest.mark.gpu


def _nets(fixtures):
    yield "fixture-rnd-0", fixture_network(fixtures, "rnd", 0)
    yield "fixture-xor-3", fixture_network(fixtures, "xor", 3)
    yield "c0-like-N30-P8", random_network(0, 30, 8, M=3, I_0=50.0, a=0.1, vmax=1.0)
    yield "N12-P0", random_network(5, 12, 0, M=6, I_0=20.0, a=0.3)
    yield "N60-P10", random_network(6, 60, 10, I_0=50.0, a=0.1)


def test_replay_event_sequence_bit_exact(fixtures, gpu_engine):
    hops = 10000
    for name, net in _nets(fixtures):
        print(f"[gpu] replay {name} N={net['N']} P={net['P']} hops={hops}", flush=True)
        N, P = net["N"], net["P"]
        U = np.random.RandomState(N * 7 + P).rand(2 * hops)
        ref = oracle.numba_simulate(N, P, net["nu"], net["kT"], net["I_0"], net["R"],
                                    net["distances"], net["E_constant"], net["site_energies"],
                                    net["transitions_constant"], net["occupation"],
                                    np.zeros(P, np.int64), hops, record=True, uniforms=U, log=True)
        assert ref.status == 0
        se = net["site_energies"].copy()
        occ = net["occupation"].copy()
        eo = np.zeros(P, np.int64)
        t, traffic, oit, ev, tot = engine.simulate_discrete(
            N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
            net["E_constant"], se, net["transitions_constant"], occ, eo, hops, record=True,
            uniforms=U, log=True)
        ulp = np.abs(tot - ref.totals) / np.spacing(ref.totals)
        first_diff = np.flatnonzero(ev != ref.events)
        print(f"[gpu]   events equal: {first_diff.size == 0}; totals max ulp diff {ulp.max():.0f};"
              f" time rel diff {abs(t - ref.time) / ref.time:.2e}", flush=True)
        assert first_diff.size == 0, f"{name}: first differing hop {first_diff[:5]}"
        assert np.array_equal(eo, ref.electrode_occupation)
        assert np.array_equal(occ, ref.occupation)
        assert abs(t - ref.time) <= 1e-12 * ref.time
        assert ulp.max() <= 64
        assert np.allclose(se[:N], ref.site_energies[:N], rtol=1e-13, atol=1e-12)
        assert np.array_equal(traffic, ref.traffic)
        assert np.allclose(oit, ref.occupations_in_time, rtol=1e-12, atol=0)


def test_philox_batch_matches_oracle_per_trajectory(fixtures, gpu_engine):
    net = random_network(0, 30, 8, M=3, I_0=50.0, a=0.1, vmax=1.0)
    N, P, K, hops = 30, 8, 48, 3000
    seeds = np.arange(K, dtype=np.uint64) * 7919 + 3
    eo_ref, t_ref, occ_ref, st = oracle.numba_ensemble(
        N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"], net["E_constant"],
        net["site_energies"], net["transitions_constant"], net["occupation"], seeds, hops, hop0=5)
    assert (st == 0).all()
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    b = engine.Batch(engine.make_trajs(K, lay.id, es[0], seeds, hops, net["occupation"], hop0=5))
    b.run(precision=_lib.FP64_DENSE)
    ms = b.sync()
    r = b.results()
    print(f"[gpu] fp64 batch {K}x{hops} hops in {ms:.1f} ms", flush=True)
    assert (r.status == 0).all() and (r.hops_done == hops).all()
    assert np.array_equal(r.eo[:, :P], eo_ref)
    assert np.array_equal(engine.bits_to_occupation(r.occ_bits, N), occ_ref.astype(bool))
    assert np.allclose(r.time, t_ref, rtol=1e-11, atol=0)


def test_prehops_continue_the_stream(gpu_engine):
    net = random_network(3, 20, 6, I_0=50.0, a=0.1)
    N, P = 20, 6
    pre, hops = 150, 700
    a = oracle.numba_simulate(*[net[k] for k in ("N", "P", "nu", "kT", "I_0", "R", "distances",
                                                 "E_constant", "site_energies",
                                                 "transitions_constant", "occupation")],
                              np.zeros(P, np.int64), pre, seed=77, hop0=0)
    b_ = oracle.numba_simulate(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                               net["E_constant"], net["site_energies"], net["transitions_constant"],
                               a.occupation, np.zeros(P, np.int64), hops, seed=77, hop0=pre)
    lay = engine.Layout(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    bt = engine.Batch(engine.make_trajs(1, lay.id, es[0], 77, hops, net["occupation"], prehops=pre))
    bt.run(precision=_lib.FP64_DENSE)
    bt.sync()
    r = bt.results()
    assert np.array_equal(r.eo[0, :P], b_.electrode_occupation)
    assert abs(r.time[0] - b_.time) <= 1e-12 * b_.time


def test_zero_rate_is_reported(gpu_engine):
    net = random_network(9, 5, 0, M=0)
    se = net["site_energies"].copy()
    occ = np.ones(5, bool)
    with pytest.raises(_lib.KmcError) as e:
        engine.simulate_discrete(5, 0, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                                 net["E_constant"], se, net["transitions_constant"], occ,
                                 np.zeros(0, np.int64), 3)
    assert e.value.code == _lib.ERR_ZERO_RATE


def test_zero_event_uniform_selects_the_first_pair(gpu_engine):
    """u == 0 passes the reference's `problist[i] >= event` test at i = 0
    (kmc_dopant_networks.py:106-110): the (0, 0) self-pair, which empties and
    refills acceptor 0, so acceptor 0 ends occupied.  The FP64 kernel keeps
    that quirk hop for hop with the oracle."""
    net = random_network(9, 20, 6, I_0=30.0, a=0.3, vmax=2.0)
    N, P, hops = 20, 6, 400
    U = np.random.RandomState(3).rand(2 * hops)
    zero_hops = np.arange(5, hops, 37)
    U[2 * zero_hops + 1] = 0.0
    ref = oracle.numba_simulate(N, P, net["nu"], net["kT"], net["I_0"], net["R"],
                                net["distances"], net["E_constant"], net["site_energies"],
                                net["transitions_constant"], net["occupation"],
                                np.zeros(P, np.int64), hops, uniforms=U, log=True)
    se = net["site_energies"].copy()
    occ = net["occupation"].copy()
    eo = np.zeros(P, np.int64)
    print("[gpu] u == 0 hops on the FP64 kernel", flush=True)
    t, _, _, ev, _ = engine.simulate_discrete(
        N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"], net["E_constant"], se,
        net["transitions_constant"], occ, eo, hops, uniforms=U, log=True)
    assert (ev[zero_hops] == 0).all() and (ref.events[zero_hops] == 0).all()
    assert np.array_equal(ev, ref.events)
    assert np.array_equal(occ, ref.occupation) and np.array_equal(eo, ref.electrode_occupation)
