"""kmc_dn host mirror (set-up only; no device needed) against the reference fixtures."""
import numpy as np

from kmc_dn_amd.kmc import kmc_dn


def test_constructor_reproduces_fixture_tables(fixtures):
    k = 7
    acc = fixtures["xor_acceptors"][k]
    don = fixtures["xor_donors"][k]
    el = fixtures["xor_electrodes"][k]
    m = kmc_dn(30, 3, 1, 1, 0, mu=0, I_0=100, a=0.25, electrodes=el, acceptors=acc, donors=don,
               res=0.01)
    assert m.R == fixtures["R"] and m.ab == fixtures["ab"] and m.dim == 2 and m.P == 8
    assert np.max(np.abs(m.eV_constant - fixtures["xor_eV_constant"][k])) < 1e-10
    assert np.allclose(m.comp_constant, fixtures["xor_comp_constant"][k], rtol=4e-16, atol=0)
    assert np.array_equal(m.site_energies[30:], el[:, 3])
    assert m.transitions_constant.shape == (38, 38)
    # loaded dopants: charges are not placed (kmc_dopant_networks.py:404-405)
    assert m.occupation.sum() == 0


def test_random_placement_uses_numpy_global_generator():
    np.random.seed(3)
    a = kmc_dn(10, 2, 1, 1, 0, electrodes=np.zeros((0, 4)))
    np.random.seed(3)
    b = kmc_dn(10, 2, 1, 1, 0, electrodes=np.zeros((0, 4)))
    assert np.array_equal(a.acceptors, b.acceptors) and np.array_equal(a.occupation, b.occupation)
    assert a.occupation.sum() == 8


def test_save_load_round_trip(tmp_path, fixtures):
    el = fixtures["rnd_electrodes"][0]
    m = kmc_dn(30, 3, 1, 1, 0, I_0=100, electrodes=el, acceptors=fixtures["rnd_acceptors"][0],
               donors=fixtures["rnd_donors"][0])
    m.current = np.arange(8.0)
    p = tmp_path / "m.kmc"
    m.saveSelf(str(p))
    n = kmc_dn(30, 3, 1, 1, 0, I_0=100, electrodes=el, acceptors=fixtures["rnd_acceptors"][0],
               donors=fixtures["rnd_donors"][0])
    n.loadSelf(str(p))
    assert np.array_equal(n.expected_current, np.arange(8.0))
    assert np.array_equal(n.E_constant, m.E_constant)
