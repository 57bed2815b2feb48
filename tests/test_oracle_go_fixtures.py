"""The oracle's Go-loop restatement against the Go library's own output.

thesis_indrek/tests/{rnd_min_max5M, XOR_wide5M} hold, per fixture network, the
mean and population std of 5 runs of wrapperSimulateRecordPlus at 5e6 hops
(thesis_indrek/generate_tests.py:44-51: go_simulation from an empty network).
The restatement (oracle_go_simulate, simulation.go:327-432 with its occupation
cache) runs the same protocol on 8 of those networks; its currents agree with
the Go library's within the fixtures' own 5-run scatter, with the agreement
statistic the reference's engines show among themselves (median |z| 0.74,
95th percentile 2.6), and its E_DB (validate_tests.py:80-135) is within the
range the reference published for its Go back end at 5e6 hops.
"""
import numpy as np

import oracle
from tests.netgen import fixture_network
from tests.stats import e_db, mean_std_pop


def test_go_restatement_reproduces_the_go_made_fixtures(fixtures):
    runs, hops = 5, 5_000_000
    zs, new_m, new_s, ref_m, ref_s = [], [], [], [], []
    for lset, sname, ks in [("rnd", "rnd_min_max5M", (0, 31, 62, 84)),
                            ("xor", "XOR_wide5M", (5, 47, 71, 93))]:
        for k in ks:
            net = fixture_network(fixtures, lset, k)
            N, P = net["N"], net["P"]
            eo, tm, _, st = oracle.go_ensemble(
                N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                net["E_constant"], net["site_energies"], net["transitions_constant"],
                np.zeros(N, np.uint8), np.arange(1000 * k, 1000 * k + runs, dtype=np.uint64),
                hops, full_recompute=True, use_cache=True, cache_max_bytes=1 << 27)
            assert (st == 0).all()
            m, s = mean_std_pop(eo / tm[:, None])
            rm, rs = fixtures[f"{sname}_mean_currents"][k], fixtures[f"{sname}_stddev_currents"][k]
            se = np.sqrt(s ** 2 / (runs - 1) + rs ** 2 / (runs - 1))
            big = np.abs(rm) > 1e-5
            zs.append(np.abs(m - rm)[big] / se[big])
            new_m.append(m), new_s.append(s), ref_m.append(rm), ref_s.append(rs)
    z = np.concatenate(zs)
    e, sd, _ = e_db(np.array(ref_m), np.array(ref_s), np.array(new_m), np.array(new_s))
    print(f"Go restatement vs Go-made 5e6 fixtures: {z.size} currents, median |z| "
          f"{np.median(z):.2f}, p95 {np.percentile(z, 95):.2f}, E_DB {e:.3f} (sigma {sd:.3f})")
    assert z.size >= 25
    assert np.median(z) < 1.2 and np.percentile(z, 95) < 3.5
    assert e <= 0.47   # the reference's Go back end at 5e6 hops: 0.17-0.47 (printTables.py:119,136)
