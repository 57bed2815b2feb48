"""P1 Laplace restatement (kmc_dopant_networks.py:706-821) against the reference's
stored eV_constant, plus linearity of the unit-voltage basis."""
import numpy as np
import pytest

from kmc_dn_amd import laplace


@pytest.mark.parametrize("lset", ["rnd", "xor"])
def test_eV_matches_fixtures(fixtures, lset):
    worst = 0.0
    for k in range(0, 100, 3):
        acc = fixtures[f"{lset}_acceptors"][k]
        el = fixtures[f"{lset}_electrodes"][k]
        ev = laplace.eV_constant(acc, el, 1.0, 1.0, 0.0, res=float(fixtures["res"]),
                                 mu=float(fixtures["mu"]))
        worst = max(worst, np.max(np.abs(ev - fixtures[f"{lset}_eV_constant"][k])))
    # FEniCS' direct solve vs SuperLU on the same P1 system: ~1e-11 absolute on |V| <= 300
    assert worst < 1e-10, worst


def test_mesh_matches_reference_resolution():
    lap = laplace.Laplace2D(1.0, 1.0, 0.01)
    assert lap.mesh.nx == 99 and lap.mesh.ny == 99      # int(1 // 0.01) == 99
    assert lap.mesh.xs[-1] == 1.0


def test_superposition_equals_direct_solve(fixtures):
    acc = fixtures["rnd_acceptors"][5]
    el = fixtures["rnd_electrodes"][5].copy()
    basis = laplace.LaplaceBasis(acc, el, 1.0, 1.0, res=0.01)
    rng = np.random.RandomState(0)
    for _ in range(8):
        V = rng.uniform(-2, 2, 8)
        mu = rng.uniform(-1, 1)
        el[:, 3] = V
        direct = laplace.eV_constant(acc, el, 1.0, 1.0, 0.0, res=0.01, mu=mu)
        assert np.max(np.abs(basis.eV(V, mu) - direct)) < 1e-10


def test_constant_boundary_gives_constant_potential():
    acc = np.random.RandomState(1).rand(10, 3) * [1, 1, 0]
    el = np.array([[0, 0.5, 0, 3.0], [1, 0.5, 0, 3.0], [0.5, 0, 0, 3.0], [0.5, 1, 0, 3.0]])
    ev = laplace.eV_constant(acc, el, 1.0, 1.0, 0.0, res=0.05, mu=3.0)
    assert np.allclose(ev, 3.0, atol=1e-12)


def test_1d_is_linear_between_electrodes():
    acc = np.array([[0.25, 0, 0], [0.5, 0, 0], [0.8, 0, 0]])
    el = np.array([[0, 0, 0, 2.0], [1, 0, 0, -1.0]])
    ev = laplace.eV_constant(acc, el, 1.0, 0.0, 0.0)
    assert np.allclose(ev, 2.0 - 3.0 * acc[:, 0], atol=1e-12)
