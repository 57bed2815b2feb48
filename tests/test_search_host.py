"""Search error functions (dn_search.py:98-148, voltage_search.py:88-131) on
hand-computed inputs, and the per-test voltage vectors."""
import numpy as np
import pytest

from kmc_dn_amd import search

XOR = [((0, 0), False), ((1, 0), True), ((0, 1), True), ((1, 1), False)]


def test_average_cumulative_error():
    # only 0.05 exceeds 0.02, by 0.03, over three differences
    assert search.average_cumulative_error([0.01, 0.05, 0.002], 0.02) == pytest.approx(0.01)
    assert search.average_cumulative_error([0.01, 0.002], 0.02) == 0.0
    with pytest.raises(ValueError):
        search.average_cumulative_error([], 0.02)


def test_dn_search_error():
    currents = np.array([[1.0, 2.0, 3.0], [0.5, 0.0, -1.0]])
    tests = [((0.1, 0.2), [(0, 1.1), (2, 3.0)]), ((0.0, 0.3), [(2, -0.5)])]
    # |1.0 - 1.1| = 0.1, |3 - 3| = 0, |-1 - (-0.5)| = 0.5; excess over 0.05: 0.05 + 0.45
    assert search.dn_search_error(currents, tests, 0.05) == pytest.approx(0.5 / 3)
    assert search.dn_search_error(currents, tests, 0.05, use_tests=1) == pytest.approx(0.025)


def test_corr_error_overlapping_classes():
    # true outputs 5, 6 above false outputs 1, 2 (x 1e-3): separation 2e-3 - 5e-3;
    # centred pattern (-5, 5, 5, -5), centred values (-2.5, 1.5, 2.5, -1.5):
    # corr = 40 / sqrt(100 * 17)
    v = np.array([1.0, 5.0, 6.0, 2.0]) * 1e-3
    corr = 40.0 / np.sqrt(1700.0)
    assert search.corr_error(v, XOR) == pytest.approx(corr * -3e-3, rel=1e-12)
    assert search.corr_error(v, XOR, corr_pow=2) == pytest.approx(corr ** 2 * -3e-3, rel=1e-12)


def test_corr_error_separated_classes_return_the_separation():
    v = np.array([5.0, 1.0, 2.0, 6.0]) * 1e-3      # a false test above every true one
    assert search.corr_error(v, XOR) == pytest.approx(6e-3 - 1e-3)
    assert search.diff_error(v, XOR) == pytest.approx(5e-3)


def test_voltages_per_test():
    V = np.arange(8, dtype=float)
    out = search.voltages_per_test(V, XOR)
    assert out.shape == (4, 8)
    assert np.array_equal(out[:, :2], np.array([t[0] for t in XOR], float))
    assert (out[:, 2:] == V[2:]).all()
    d = search.voltages_per_test(V, XOR, mode="diff", true_voltage=0.7)
    assert np.array_equal(d[:, :2], 0.7 * np.array([t[0] for t in XOR], float))
