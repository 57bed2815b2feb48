"""Constant tables (kmc_dn.initialize, kmc_dopant_networks.py:408-452) against the
reference's own stored tables in its .kmc fixtures."""
import numpy as np
import pytest

from kmc_dn_amd import tables as T


def test_R_and_ab(fixtures):
    dim, R = T.characteristic_length(int(fixtures["N"]), 1.0, 1.0, 0.0)
    assert dim == 2 and R == fixtures["R"]
    assert 0.25 * R == fixtures["ab"]


@pytest.mark.parametrize("lset", ["rnd", "xor"])
def test_distances_and_tc(fixtures, lset):
    d = T.calc_distances(fixtures[f"{lset}_acceptors"][0], fixtures[f"{lset}_electrodes"][0])
    assert np.array_equal(d, fixtures[f"{lset}0_distances"])
    tc = T.calc_transitions_constant(d, float(fixtures["ab"]), 1.0)
    ref = fixtures[f"{lset}0_transitions_constant"]
    assert np.all(np.diag(tc) == 0) and np.all(np.diag(ref) == 0)
    # numpy's vectorised exp may differ from the fixture writer's by 1 ulp
    assert np.max(np.abs(tc - ref)) <= 1.2e-16
    assert np.max(np.abs(tc - ref) / np.maximum(np.abs(ref), 1e-300)) < 4.5e-16


@pytest.mark.parametrize("lset", ["rnd", "xor"])
def test_comp_and_E_constant(fixtures, lset):
    I0R = float(fixtures["I_0"]) * float(fixtures["R"])
    for k in range(100):
        comp = T.calc_comp_constant(fixtures[f"{lset}_acceptors"][k], fixtures[f"{lset}_donors"][k],
                                    float(fixtures["I_0"]), float(fixtures["R"]))
        ref = fixtures[f"{lset}_comp_constant"][k]
        # the fixtures were written by a dist() that summed with np.sum; the
        # current source (kmc_dopant_networks.py:1003-1006) differs by <= 1 ulp
        assert np.all(np.abs(comp - ref) <= 2 * np.spacing(np.abs(ref))), k
        E = T.calc_E_constant(fixtures[f"{lset}_eV_constant"][k], ref)
        assert np.array_equal(E, fixtures[f"{lset}_E_constant"][k])
    assert I0R > 0


def test_place_charges_counts():
    rng = np.random.RandomState(0)
    occ = T.place_charges_random(rng, 30, 3)
    assert occ.sum() == 27 and occ.dtype == bool


def test_total_energy_matches_closed_form():
    rng = np.random.RandomState(1)
    acc = rng.rand(6, 3) * [1, 1, 0]
    d = T.calc_distances(acc, np.zeros((0, 4)))
    eV = rng.randn(6)
    occ = np.array([1, 0, 1, 1, 0, 0], bool)
    e = 1 - occ.astype(float)
    brute = 5.0 * sum(e[i] * e[j] / d[i, j] for i in range(6) for j in range(i + 1, 6)) - e @ eV
    assert abs(T.total_energy(occ, d, eV, 5.0, 1.0) - brute) < 1e-12
