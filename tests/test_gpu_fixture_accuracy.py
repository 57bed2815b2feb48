"""Accuracy on the reference's own fixture networks (thesis_indrek/tests).

* The 1 % bar at 1e6 hops: FP32 against the FP64 path (the numba loop's
  arithmetic) on 10 fixture networks x 256 trajectories.  Every current lies
  within 3 sigma; every current whose 3-sigma band is inside 1 % (|I| > 300
  standard errors, where a 1 % bar can be tested at all) lies within 1 %.
* The reference's own accuracy statistic, E_DB (the mean Bhattacharyya
  distance of 5-run current estimates, thesis_indrek/validate_tests.py:80-135),
  for go_simulation's back end against the Go-made 5e6-hop fixture sets, and
  for the pruned pair sets of wrapperSimulatePruned (simulation.go:199-215)
  against the 1e6-hop sets, beside the figures the reference published for
  its own engines (thesis_indrek/printTables.py:116-136).
"""
import numpy as np
import pytest

from kmc_dn_amd import _lib, engine, validate
from tests.netgen import fixture_network
from tests.stats import assert_within, compare

pytestmark = pytest.mark.gpu


def test_fp32_within_one_percent_of_fp64_at_1e6_hops(gpu_engine):
    print("[gpu] 10 fixtures x 256 trajectories x 1e6 hops, FP32 and FP64", flush=True)
    r = validate.fixtures_check(10, 256, 1_000_000)
    print(f"[gpu] {r}", flush=True)
    assert r["currents_compared"] >= 60 and r["currents_1pct_testable"] >= 30
    assert r["max_rel_1pct_testable"] < 0.01
    assert r["max_z"] < 3.0


def _edb(sets, **kw):
    cur = validate.edb_runs(sets, **kw)
    return validate.edb_table(sets, cur)


def test_edb_against_go_made_5e6_fixtures(gpu_engine):
    """go_simulation (empty start, FP32) at 5e6 hops against the sets the Go
    library made with wrapperSimulateRecordPlus at 5e6 hops (generate_tests.py:44-51):
    the reference's own 5e6 rows score 0.17-0.47."""
    print("[gpu] E_DB 2 x 100 fixtures x 5 runs x 5e6 hops", flush=True)
    t = _edb(["rnd_min_max5M", "XOR_wide5M"], hops=5_000_000)
    print(f"[gpu] E_DB vs Go-made 5e6 sets: {t}", flush=True)
    ref = validate.REFERENCE_EDB["Go recording 5e6"]
    assert t["rnd_min_max5M"][0] <= max(ref["RND"])
    assert t["XOR_wide5M"][0] <= max(ref["XOR"])


def test_edb_of_pruned_pair_sets(gpu_engine):
    """Pruning at 1e-5 of the largest tc drops pairs that carry current: the
    reference measured E_DB 58-340 there, against 0.4-3.1 at 1e-7 and 1e-9.
    The same pair sets here give the same split."""
    out = {}
    for pr in (1e-5, 1e-7, 1e-9):
        print(f"[gpu] E_DB pruned {pr:g}, 2 x 100 fixtures x 5 runs x 1e6 hops", flush=True)
        out[pr] = _edb(["rnd_min_max", "XOR_wide"], hops=1_000_000, prune=pr)
    unpruned = _edb(["rnd_min_max", "XOR_wide"], hops=1_000_000)
    print(f"[gpu] E_DB pruned: {out}; unpruned {unpruned}", flush=True)
    for name, fam in [("rnd_min_max", "RND"), ("XOR_wide", "XOR")]:
        assert out[1e-5][name][0] > 10
        for pr, row in [(1e-7, "Go pruning 1e-7"), (1e-9, "Go pruning 1e-9")]:
            assert out[pr][name][0] <= max(validate.REFERENCE_EDB[row][fam])
        assert unpruned[name][0] <= max(validate.REFERENCE_EDB["Go 1e6"][fam])


@pytest.mark.parametrize("prune", [1e-5, 1e-7])
def test_pruned_currents_against_the_go_restatement(fixtures, gpu_engine, prune):
    """FP32 kernel on a pruned layout vs the oracle's Go loop with the same
    prune_threshold (simulation.go:199-215), same start, 3 sigma per current."""
    import oracle
    K, hops = 192, 20_000
    for lset, k in [("rnd", 3), ("xor", 7)]:
        net = fixture_network(fixtures, lset, k)
        N, P = net["N"], net["P"]
        lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                            net["transitions_constant"], prune)
        es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])[0]
        empty = np.zeros(N, bool)
        b = engine.Batch(engine.make_trajs(K, lay.id, es, np.arange(K, dtype=np.uint64), hops,
                                           empty))
        print(f"[gpu] pruned {prune:g} fixture {lset}{k}: FP32 {K} x {hops}", flush=True)
        b.run(precision=_lib.FP32_BATCHED)
        b.sync()
        r = b.results()
        b.close()
        lay.close()
        assert (r.status == 0).all()
        eo, tm, _, st = oracle.go_ensemble(
            N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"], net["E_constant"],
            net["site_energies"], net["transitions_constant"], empty.astype(np.uint8),
            np.arange(10_000, 10_000 + K, dtype=np.uint64), hops, use_cache=True,
            prune_threshold=prune)
        assert (st == 0).all()
        c = compare(r.currents(P), eo / tm[:, None])
        assert_within(c, f"pruned {prune:g} {lset}{k}: FP32 GPU vs Go restatement")
