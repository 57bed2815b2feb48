"""Limits and edge cases of the batch C-ABI: largest networks, clean errors past
the kernels' limits, empty budgets, networks without electrodes, and a
trajectory that stops on a zero total rate without disturbing the others."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, engine
from kmc_dn_amd import tables as T
from tests.netgen import random_network

pytestmark = pytest.mark.gpu


def _layout(net, positions=True):
    N, P = net["N"], net["P"]
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    if positions:
        lay.set_positions(0.1 * net["R"], T.site_positions(net["acceptors"], net["electrodes"]))
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    return lay, es[0]


def _invariants(net, trajs, r, hops):
    N, P = net["N"], net["P"]
    assert (r.status == 0).all() and (r.hops_done == hops).all() and (r.time > 0).all()
    n0 = engine.bits_to_occupation(trajs["occ"], N).sum(1)
    n1 = engine.bits_to_occupation(r.occ_bits, N).sum(1)
    assert np.array_equal(r.eo[:, :P].sum(1), n0 - n1)


def test_largest_fp32_network(gpu_engine):
    """S = 192 (N = 176, P = 16): three site chunks, positions form and team."""
    net = random_network(11, 176, 16, I_0=50.0, a=0.1, vmax=1.0)
    lay, es = _layout(net)
    K, hops = 64, 3000
    trajs = engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                              net["occupation"])
    b = engine.Batch(trajs)
    try:
        for team in (1, 4):
            print(f"[gpu] S=192 positions team={team}", flush=True)
            b.run(precision=_lib.FP32_BATCHED, table_form=3, team=team)
            b.sync()
            _invariants(net, trajs, b.results(), hops)
        # a {lt, lt+c} table of S = 192 does not fit 160 KiB of LDS: a clean error
        with pytest.raises(_lib.KmcError) as e:
            b.run(precision=_lib.FP32_BATCHED, table_form=1)
        assert e.value.code == _lib.ERR_LIMIT
        # nor does the FP64 dense kernel take S > 80
        with pytest.raises(_lib.KmcError) as e:
            b.run(precision=_lib.FP64_DENSE)
        assert e.value.code == _lib.ERR_LIMIT
    finally:
        b.close()
        lay.close()


def test_network_above_the_fp32_limit_is_refused(gpu_engine):
    net = random_network(12, 180, 16, I_0=50.0, a=0.1)     # S = 196 > 192 sites
    with pytest.raises(_lib.KmcError) as e:
        engine.Layout(180, 16, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                      net["transitions_constant"])
    assert e.value.code == _lib.ERR_LIMIT


def test_fp64_at_its_size_limit(gpu_engine):
    net = random_network(13, 72, 8, I_0=50.0, a=0.1)       # S = 80
    lay, es = _layout(net, positions=False)
    K, hops = 8, 500
    trajs = engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                              net["occupation"])
    b = engine.Batch(trajs)
    try:
        print("[gpu] FP64 dense at S=80", flush=True)
        b.run(precision=_lib.FP64_DENSE)
        b.sync()
        _invariants(net, trajs, b.results(), hops)
    finally:
        b.close()
        lay.close()


def test_zero_budget_and_no_electrodes(gpu_engine):
    """hops = 0 leaves the state as it was; P = 0 conserves the number of holes."""
    net = random_network(14, 40, 0, M=4, I_0=50.0, a=0.1)
    lay, es = _layout(net)
    K = 32
    hops = np.where(np.arange(K) < 4, 0, 2000)
    trajs = engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                              net["occupation"])
    b = engine.Batch(trajs)
    try:
        for prec, form in ((_lib.FP32_BATCHED, 1), (_lib.FP32_BATCHED, 3), (_lib.FP64_DENSE, 0)):
            print(f"[gpu] P=0 precision={prec} form={form}", flush=True)
            b.run(precision=prec, table_form=form)
            b.sync()
            r = b.results()
            assert (r.status == 0).all()
            assert (r.hops_done == hops).all()
            assert (r.time[:4] == 0).all() and (r.time[4:] > 0).all()
            assert np.array_equal(r.occ_bits[:4], trajs["occ"][:4])
            n0 = net["occupation"].sum()
            assert (engine.bits_to_occupation(r.occ_bits, 40).sum(1) == n0).all()
    finally:
        b.close()
        lay.close()


def test_zero_rate_trajectory_stops_alone(gpu_engine):
    """Every acceptor empty and no electrode: no possible hop.  That trajectory
    reports ZERO_RATE; trajectories of a healthy layout in the same launch run on."""
    dead = random_network(15, 20, 0, M=0, I_0=50.0, a=0.1)
    dead["occupation"] = np.zeros(20, bool)
    good = random_network(16, 30, 8, I_0=50.0, a=0.1)
    ld, ed = _layout(dead)
    lg, eg = _layout(good)
    t = np.concatenate([
        engine.make_trajs(4, ld.id, ed, np.arange(4, dtype=np.uint64), 1000, dead["occupation"]),
        engine.make_trajs(16, lg.id, eg, np.arange(16, dtype=np.uint64), 1000, good["occupation"])])
    b = engine.Batch(t)
    try:
        for form in (3,):
            print(f"[gpu] zero-rate trajectories in a mixed batch, form {form}", flush=True)
            b.run(precision=_lib.FP32_BATCHED, table_form=form)
            b.sync()
            r = b.results()
            assert (r.status[:4] == _lib.ERR_ZERO_RATE).all() and (r.hops_done[:4] == 0).all()
            assert (r.status[4:] == 0).all() and (r.hops_done[4:] == 1000).all()
    finally:
        b.close()
        ld.close()
        lg.close()
