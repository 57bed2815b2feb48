"""Batch state across runs: record buffers after updates, graphs after the
registry moves, traffic inside a graph, and stale results refused."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, engine
from tests.netgen import random_network

pytestmark = pytest.mark.gpu


def _layout(net):
    N, P = net["N"], net["P"]
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    return lay, es[0]


def _trajs(lay, es, net, K, hops, seed0=0):
    return engine.make_trajs(K, lay.id, es, np.arange(seed0, seed0 + K, dtype=np.uint64), hops,
                             net["occupation"])


def _run_all(b, **kw):
    b.run(precision=_lib.FP32_BATCHED, record=True, traffic=True, **kw)
    b.sync()
    return b.results(), b.traffic_flat()


def test_update_to_larger_network_and_more_trajectories(gpu_engine):
    small = random_network(31, 12, 4, I_0=30.0, a=0.2, vmax=2.0)
    large = random_network(32, 50, 10, I_0=30.0, a=0.2, vmax=2.0)
    ls, es_s = _layout(small)
    ll, es_l = _layout(large)
    print("[gpu] batch update: S 16 -> 60, 8 -> 40 trajectories, record + traffic", flush=True)
    b = engine.Batch(_trajs(ls, es_s, small, 8, 3000))
    _run_all(b)
    t_large = _trajs(ll, es_l, large, 40, 3000, seed0=100)
    b.update(t_large)
    r_upd, tr_upd = _run_all(b)
    fresh = engine.Batch(t_large)
    r_new, tr_new = _run_all(fresh)
    assert np.array_equal(r_upd.eo, r_new.eo)
    assert np.array_equal(r_upd.time, r_new.time)
    assert np.array_equal(r_upd.occ_time, r_new.occ_time)
    assert np.array_equal(tr_upd, tr_new)
    S = large["N"] + large["P"]
    assert tr_upd.shape[1] == S * S and np.abs(tr_upd).sum() > 0
    # back to fewer trajectories in the same buffers
    b.update(t_large[:5])
    r5, tr5 = _run_all(b)
    assert np.array_equal(r5.eo, r_new.eo[:5]) and np.array_equal(tr5, tr_new[:5])
    b.close()
    fresh.close()
    ls.close()
    ll.close()


def test_graph_replays_after_the_registry_grows(gpu_engine):
    net = random_network(33, 30, 8, I_0=50.0, a=0.1, vmax=1.0)
    lay, es = _layout(net)
    b = engine.Batch(_trajs(lay, es, net, 64, 2000))
    b.run(precision=_lib.FP32_BATCHED)
    b.sync()
    eager = b.results()
    b.run(precision=_lib.FP32_BATCHED, use_graph=True)
    b.sync()
    assert np.array_equal(b.results().eo, eager.eo)
    extra = []
    for k in range(6):
        # many layouts and energy sets: the registry's device arrays move
        other = random_network(40 + k, 20, 6, I_0=50.0, a=0.1, vmax=1.0)
        lo, _ = _layout(other)
        lo.energy_sets(np.repeat(other["E_constant"][None], 300, 0),
                       np.repeat(other["site_energies"][None, 20:], 300, 0))
        extra.append(lo)
        print(f"[gpu] graph replay after registry growth {k}", flush=True)
        b.run(precision=_lib.FP32_BATCHED, use_graph=True)
        b.sync()
        r = b.results()
        assert np.array_equal(r.eo, eager.eo) and np.array_equal(r.time, eager.time)
    for lo in extra:
        lo.close()
    b.run(precision=_lib.FP32_BATCHED, use_graph=True)
    b.sync()
    assert np.array_equal(b.results().eo, eager.eo)
    b.close()
    lay.close()


def test_traffic_inside_a_graph_equals_eager(gpu_engine):
    net = random_network(34, 24, 6, I_0=20.0, a=0.3, vmax=2.0)
    lay, es = _layout(net)
    b = engine.Batch(_trajs(lay, es, net, 32, 3000))
    r0, tr0 = _run_all(b)
    for rep in range(3):
        print(f"[gpu] graph replay {rep} with traffic", flush=True)
        r, tr = _run_all(b, use_graph=True)
        assert np.array_equal(tr, tr0) and np.array_equal(r.eo, r0.eo)
        b.run(precision=_lib.FP32_BATCHED)   # eager work between replays
        b.sync()
    b.close()
    lay.close()


def test_stale_records_are_refused(gpu_engine):
    net = random_network(35, 20, 4, I_0=20.0, a=0.3, vmax=2.0)
    lay, es = _layout(net)
    b = engine.Batch(_trajs(lay, es, net, 4, 500))
    _run_all(b)
    b.run(precision=_lib.FP32_BATCHED)
    b.sync()
    with pytest.raises(_lib.KmcError):
        b.traffic_flat()
    b.run(precision=_lib.FP64_DENSE)
    b.sync()
    with pytest.raises(_lib.KmcError):
        b.traffic_flat()
    with pytest.raises(_lib.KmcError):
        b.run(precision=_lib.FP64_DENSE, traffic=True)
    b.close()
    lay.close()
