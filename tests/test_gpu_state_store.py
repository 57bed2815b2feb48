"""The LDS state store of one-chunk solo trajectories changes no result.

A new state found among the states a trajectory's two register entries evicted
is swapped back into the entry instead of evaluated.  Between two FP64
recomputes the energies of a state are a function of its occupation alone (the
recompute's energies plus the Coulomb columns of the acceptors changed since,
in ascending site order), so the restored entry holds the bits an evaluation
gives: runs with the store off (state_store = -1), with 1 or 2 slots and with
the automatic size (0) agree bit for bit, in every pair-term form and with
records; one-chunk networks run inside a three-chunk positions kernel (whose
one-chunk code takes no store) equal the same trajectories in a one-chunk batch
of their own.  The phase probe counts restores only with the store on.  No reference counterpart
(the Go library's occupation cache, simulation.go:254-296, keeps every state).
"""
import ctypes as C

import numpy as np
import pytest

from kmc_dn_amd import _lib, engine, tables, workloads
from tests.netgen import random_network

pytestmark = pytest.mark.gpu


def _bits(r, rec=False):
    out = [r.time.view(np.uint64).copy(), r.eo.copy(), r.occ_bits.copy(), r.status.copy(),
           r.hops_done.copy()]
    if rec:
        out.append(r.occ_time.view(np.uint64).copy())
    return out


def _c1_layout():
    net = workloads.network_2d(0)
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
    es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])
    return net, lay, es[0]


@pytest.mark.parametrize("form", [1, 2, 3])
def test_store_changes_no_bits(gpu_engine, form):
    net, lay, es = _c1_layout()
    K, hops = 512, 20_000
    print(f"[gpu] state store on/off, N=30 network, form {form}: {K} x {hops}", flush=True)
    try:
        b = engine.Batch(engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                                           net.occupation))
        ref = None
        for store in (-1, 0, 1, 2):
            b.run(precision=_lib.FP32_BATCHED, table_form=form, waves_per_group=4,
                  state_store=store)
            b.sync()
            r = b.results()
            assert (r.status == 0).all() and (r.hops_done == hops).all()
            if ref is None:
                ref = _bits(r)
            else:
                for x, y in zip(ref, _bits(r)):
                    assert np.array_equal(x, y), store
        # recording kernels take the store too
        recs = []
        for store in (-1, 0):
            b.run(precision=_lib.FP32_BATCHED, table_form=form, waves_per_group=4,
                  state_store=store, record=True)
            b.sync()
            recs.append(_bits(b.results(), rec=True))
        for x, y in zip(*recs):
            assert np.array_equal(x, y)
        b.close()
    finally:
        lay.close()


def test_one_chunk_code_of_a_three_chunk_kernel_equals_a_one_chunk_batch(gpu_engine):
    """A batch with an S = 140 network runs the positions kernel of three site
    chunks; its S = 38 trajectories take that kernel's one-chunk code, which the
    host gives no store (state_store 0 and -1 run alike there), and equal, bit for
    bit, the same trajectories in a batch of their own, which runs the one-chunk
    positions kernel with the store the host sizes for it."""
    small = random_network(5, 30, 8, I_0=50.0, a=0.1, vmax=1.0)
    big = random_network(6, 128, 12, I_0=50.0, a=0.1, vmax=1.0)
    lays, trajs = [], []
    for net, K in ((small, 128), (big, 16)):
        N, P = net["N"], net["P"]
        lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                            net["transitions_constant"])
        lay.set_positions(0.1 * net["R"], tables.site_positions(net["acceptors"], net["electrodes"]))
        es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
        lays.append(lay)
        trajs.append(engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), 10_000,
                                       net["occupation"]))
    try:
        print("[gpu] one-chunk trajectories in a three-chunk positions kernel vs alone", flush=True)
        mixed = engine.Batch(np.concatenate(trajs))
        outs = []
        for store in (-1, 0):   # the option is accepted; no slots in this kernel
            mixed.run(precision=_lib.FP32_BATCHED, table_form=3, state_store=store)
            mixed.sync()
            outs.append(_bits(mixed.results()))
        for x, y in zip(*outs):
            assert np.array_equal(x, y)
        mixed.close()
        alone = engine.Batch(trajs[0])
        alone.run(precision=_lib.FP32_BATCHED, table_form=3)
        alone.sync()
        a = _bits(alone.results())
        alone.close()
        for x, y in zip(a, outs[1]):
            assert np.array_equal(x, y[:128])
    finally:
        for lay in lays:
            lay.close()


def test_phase_probe_counts_restores_only_with_the_store(gpu_engine):
    net, lay, es = _c1_layout()
    K, hops = 256, 50_000
    try:
        b = engine.Batch(engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                                           net.occupation))
        got = {}
        for store in (-1, 0):
            o = _lib.kmc_run_opts(_lib.FP32_BATCHED, 4, 0, 0, 0, 0, 0, 1, 0, 1, 0, store)
            cyc = (C.c_uint64 * 11)()
            _lib.check(_lib.load().kmc_probe_fp32_phases(b.id, C.byref(o), cyc, 11))
            got[store] = list(cyc)
        print(f"[gpu] restores per hop: {got[0][10] / (K * hops):.5f}", flush=True)
        assert got[-1][10] == 0 and got[0][10] > 0
        # restored states were misses of the two entries: fewer evaluations
        assert got[0][7] == got[-1][7]      # hops from a cached register entry: same count
        b.close()
    finally:
        lay.close()


def test_store_option_is_validated(gpu_engine):
    net, lay, es = _c1_layout()
    try:
        b = engine.Batch(engine.make_trajs(8, lay.id, es, np.arange(8, dtype=np.uint64), 1000,
                                           net.occupation))
        for bad in (-2, 5):
            with pytest.raises(_lib.KmcError) as e:
                b.run(precision=_lib.FP32_BATCHED, state_store=bad)
            assert e.value.code == _lib.ERR_ARGS
        b.run(precision=_lib.FP32_BATCHED, state_store=3)   # at most 3 slots: accepted
        b.sync()
        assert (b.results().status == 0).all()
        b.close()
    finally:
        lay.close()
