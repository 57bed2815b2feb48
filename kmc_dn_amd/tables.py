"""Constant tables of a dopant network (host side, FP64, numpy).

These are the position-dependent inputs the hop loop consumes; the reference
builds them once per layout in ``kmc_dn.initialize`` (kmc_dopant_networks.py:408-452):

* ``R``, ``ab``                      -- :349-361
* ``distances`` (S x S)              -- calc_distances, :657-695 (dist, :1003-1006)
* ``transitions_constant`` (S x S)   -- calc_transitions_constant, :824-830
* ``comp_constant``/``E_constant``   -- calc_E_constant_V_comp, :865-899
* random placement                   -- place_dopants_random :621-639,
                                        place_charges_random :641-655

Site order is acceptors 0..N-1 followed by electrodes N..N+P-1, as in the
reference.  Every function here reproduces the reference's operation order so
the tables agree with the ones stored in the reference's ``.kmc`` fixtures
(checked in tests/test_tables.py).
"""
from __future__ import annotations

import math

import numpy as np


def characteristic_length(N: int, xdim: float, ydim: float, zdim: float) -> tuple[int, float]:
    """(dim, R) with R = (N / volume)^(-1/dim), kmc_dopant_networks.py:349-358."""
    if ydim == 0 and zdim == 0:
        return 1, (N / xdim) ** (-1)
    if zdim == 0:
        return 2, (N / (xdim * ydim)) ** (-1 / 2)
    return 3, (N / (xdim * ydim * zdim)) ** (-1 / 3)


def dist(ri, rj) -> float:
    """Cartesian distance in 3D, kmc_dopant_networks.py:1003-1006."""
    return math.sqrt((ri[0] - rj[0]) ** 2 + (ri[1] - rj[1]) ** 2 + (ri[2] - rj[2]) ** 2)


def site_positions(acceptors: np.ndarray, electrodes: np.ndarray) -> np.ndarray:
    """(S, 3) positions: acceptors then electrode points (electrodes[:, :3])."""
    acc = np.asarray(acceptors, dtype=np.float64).reshape(-1, 3)
    el = np.asarray(electrodes, dtype=np.float64).reshape(-1, 4)[:, :3]
    return np.concatenate([acc, el], axis=0)


def calc_distances(acceptors: np.ndarray, electrodes: np.ndarray) -> np.ndarray:
    """S x S distance matrix with a zero diagonal, kmc_dopant_networks.py:657-695.

    Vectorised, but each entry is evaluated as dist() does:
    sqrt(((dx*dx) + (dy*dy)) + (dz*dz)), every step rounded in binary64.
    """
    pos = site_positions(acceptors, electrodes)
    dx = pos[:, None, 0] - pos[None, :, 0]
    dy = pos[:, None, 1] - pos[None, :, 1]
    dz = pos[:, None, 2] - pos[None, :, 2]
    d = np.sqrt((dx * dx + dy * dy) + dz * dz)
    np.fill_diagonal(d, 0.0)
    return d


def calc_transitions_constant(distances: np.ndarray, ab: float, nu: float = 1.0) -> np.ndarray:
    """nu*exp(-2 d/ab) - I, kmc_dopant_networks.py:824-830 (diagonal nu-1 = 0 for nu=1)."""
    tc = nu * np.exp(-2 * distances / ab)
    tc -= np.eye(distances.shape[0])
    return tc


def calc_comp_constant(acceptors: np.ndarray, donors: np.ndarray, I_0: float, R: float) -> np.ndarray:
    """I_0*R*sum_k 1/dist(acceptor_i, donor_k), kmc_dopant_networks.py:893-894."""
    acc = np.asarray(acceptors, dtype=np.float64).reshape(-1, 3)
    don = np.asarray(donors, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(acc.shape[0])
    for i in range(acc.shape[0]):
        s = 0
        for k in range(don.shape[0]):
            s = s + 1 / dist(acc[i], don[k])
        out[i] += I_0 * R * s
    return out


def calc_E_constant(eV_constant: np.ndarray, comp_constant: np.ndarray | None) -> np.ndarray:
    """E = eV + comp (calc_E_constant_V_comp, :896) or E = eV (calc_E_constant_V, :860)."""
    eV = np.asarray(eV_constant, dtype=np.float64)
    return eV.copy() if comp_constant is None else eV + np.asarray(comp_constant, dtype=np.float64)


def place_dopants_random(rng: np.random.RandomState, N: int, M: int,
                         xdim: float, ydim: float, zdim: float):
    """Uniform acceptor/donor placement, kmc_dopant_networks.py:621-639."""
    acceptors = rng.rand(N, 3)
    donors = rng.rand(M, 3)
    acceptors[:, 0] *= xdim
    acceptors[:, 1] *= ydim
    acceptors[:, 2] *= zdim
    donors[:, 0] *= xdim
    donors[:, 1] *= ydim
    donors[:, 2] *= zdim
    return acceptors, donors


def place_charges_random(rng: np.random.RandomState, N: int, M: int) -> np.ndarray:
    """N-M holes by rejection sampling, kmc_dopant_networks.py:641-655."""
    occ = np.zeros(N, dtype=bool)
    placed = 0
    while placed < N - M:
        trial = rng.randint(N)
        if not occ[trial]:
            occ[trial] = True
            placed += 1
    return occ


def perimeter_electrodes(xdim: float = 1.0, ydim: float = 1.0, voltages=None) -> np.ndarray:
    """The reference's standard 8 perimeter electrodes (thesis_indrek/generate_tests.py:32-40)."""
    e = np.zeros((8, 4))
    e[0, :3] = [0, ydim / 4, 0]
    e[1, :3] = [0, 3 * ydim / 4, 0]
    e[2, :3] = [xdim, ydim / 4, 0]
    e[3, :3] = [xdim, 3 * ydim / 4, 0]
    e[4, :3] = [xdim / 4, 0, 0]
    e[5, :3] = [3 * xdim / 4, 0, 0]
    e[6, :3] = [xdim / 4, ydim, 0]
    e[7, :3] = [3 * xdim / 4, ydim, 0]
    if voltages is not None:
        e[:, 3] = voltages
    return e


def total_energy(occupation, distances, eV_constant, I_0: float, R: float) -> float:
    """Hamiltonian of kmc_dn.total_energy, kmc_dopant_networks.py:982-1001 (no comp term)."""
    N = len(occupation)
    H = 0
    for i in range(N - 1):
        for j in range(i + 1, N):
            H += (1 - occupation[i]) * (1 - occupation[j]) / distances[i, j]
    H *= I_0 * R
    for i in range(N):
        H = H - (1 - occupation[i]) * eV_constant[i]
    return H
