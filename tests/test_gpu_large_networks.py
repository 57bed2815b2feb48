"""Networks of three site chunks (S > 128) against the oracle.

The FP32 kernels with three site chunks -- the table forms up to the largest
network whose pair table fits LDS (S = 134 here) and the positions form up to
S = 192 -- in solo waves and in teams of 4, each compared with the oracle's
restatement of the Go loop (goSimulation/simulation.go:194-325) on the same
networks and start state: every per-electrode ensemble-mean current within 3
sigma.  The Go restatement runs with its occupation cache, which reuses the
rates of revisited states and changes no statistics.
"""
import numpy as np
import pytest

import oracle
from kmc_dn_amd import _lib, engine
from kmc_dn_amd import tables as T
from tests.netgen import random_network
from tests.stats import assert_within, compare

pytestmark = pytest.mark.gpu


def _oracle_currents(net, K, hops):
    N, P = net["N"], net["P"]
    eo, tm, _, st = oracle.go_ensemble(
        N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"], net["E_constant"],
        net["site_energies"], net["transitions_constant"], net["occupation"].astype(np.uint8),
        np.arange(50_000, 50_000 + K, dtype=np.uint64), hops, use_cache=True,
        cache_max_bytes=1 << 27)
    assert (st == 0).all()
    return eo / tm[:, None]


@pytest.mark.parametrize("N,P,runs", [
    (122, 12, [(2, 1), (2, 4), (3, 1), (3, 4)]),   # S = 134: {lt, c} table and positions
    (176, 16, [(3, 1), (3, 4)]),                   # S = 192: positions only
])
def test_three_chunk_networks_against_the_go_restatement(gpu_engine, N, P, runs):
    net = random_network(60 + N, N, P, I_0=30.0, a=0.2, vmax=5.0)
    K, hops = 192, 3000
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    lay.set_positions(0.2 * net["R"], T.site_positions(net["acceptors"], net["electrodes"]))
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])[0]
    b = engine.Batch(engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                                       net["occupation"]))
    ref = _oracle_currents(net, K, hops)
    for form, team in runs:
        print(f"[gpu] S={N + P} form {form} team {team}: {K} x {hops}", flush=True)
        b.run(precision=_lib.FP32_BATCHED, table_form=form, team=team)
        b.sync()
        r = b.results()
        assert (r.status == 0).all() and (r.hops_done == hops).all()
        holes = engine.bits_to_occupation(r.occ_bits, N).sum(1)
        assert np.array_equal(r.eo[:, :P].sum(1), net["occupation"].sum() - holes)
        assert_within(compare(r.currents(P), ref),
                      f"S={N + P} form {form} team {team} vs Go restatement")
    b.close()
    lay.close()
