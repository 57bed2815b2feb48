"""Host-side parts of the benchmark configurations and of the test statistics.

* C2's energy sets come from the unit-voltage Laplace basis; at 8 of its 1024
  voltages they equal a direct P1 solve of the same boundary values
  (kmc_dopant_networks.py:706-821), to 1e-10.
* The accuracy statistic of the reference's fixture harness
  (thesis_indrek/validate_tests.py:80-87) is the Bhattacharyya distance of
  two normal distributions: checked against -ln of the numerically integrated
  overlap integral int sqrt(p q) dx, and its zero-variance convention.
"""
import math

import numpy as np
from scipy import integrate

from kmc_dn_amd import laplace, workloads
from tests.stats import calc_d, e_db, mean_std_pop


def test_c2_superposed_energy_sets_equal_direct_solves():
    sw = workloads.iv_sweep(0, 1024, 64, 200_000)
    net = sw.net
    for i in np.linspace(0, 1023, 8).astype(int):
        el = net.electrodes.copy()
        el[:, 3] = sw.V_sets[i]
        direct = laplace.eV_constant(net.acceptors, el, 1.0, 1.0, 0.0, res=0.01, mu=0.0)
        E_direct = direct + net.comp_constant
        assert np.max(np.abs(sw.E_sets[i] - E_direct)) <= 1e-10
        assert sw.V_sets[i][0] == sw.voltages[i]


def test_c0_network_matches_its_description():
    net = workloads.network_2d(0)
    assert (net.N, net.M, net.P) == (30, 3, 8)
    assert net.occupation.sum() == net.N - net.M
    assert np.all(np.abs(net.electrodes[:, 3]) <= 1.0)
    assert math.isclose(net.ab, 0.1 * net.R) and net.I_0 == 50.0


def _overlap_distance(m1, s1, m2, s2):
    f = lambda x: math.sqrt(math.exp(-0.5 * ((x - m1) / s1) ** 2) / (s1 * math.sqrt(2 * math.pi))
                            * math.exp(-0.5 * ((x - m2) / s2) ** 2) / (s2 * math.sqrt(2 * math.pi)))
    lo = min(m1 - 12 * s1, m2 - 12 * s2)
    hi = max(m1 + 12 * s1, m2 + 12 * s2)
    bc, _ = integrate.quad(f, lo, hi, points=[m1, m2], limit=200, epsabs=1e-14, epsrel=1e-12)
    return -math.log(bc)


def test_bhattacharyya_distance_against_the_overlap_integral():
    rng = np.random.RandomState(3)
    for _ in range(20):
        m1, m2 = rng.uniform(-3, 3, 2)
        s1, s2 = rng.uniform(0.2, 2.5, 2)
        assert math.isclose(calc_d(m1, m2, s1, s2), _overlap_distance(m1, s1, m2, s2),
                            rel_tol=1e-7, abs_tol=1e-10)
    assert calc_d(1.0, 1.0, 0.5, 0.5) == 0.0
    assert calc_d(1.0, 2.0, 0.0, 0.5) == math.log(10000)


def test_e_db_and_population_std():
    runs = np.array([[1.0, 2.0], [3.0, 2.0], [2.0, 5.0], [2.0, 3.0], [2.0, 3.0]])
    m, s = mean_std_pop(runs)
    assert np.allclose(m, [2.0, 3.0])
    assert np.allclose(s, [math.sqrt(2 / 5), math.sqrt(6 / 5)])
    e, sd, d = e_db([1.0, 2.0], [1.0, 0.0], [1.0, 2.5], [1.0, 1.0])
    assert d[0] == 0.0 and d[1] == math.log(10000)
    assert math.isclose(e, math.log(10000) / 2) and math.isclose(sd, math.log(10000) / 2)
