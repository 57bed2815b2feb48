"""Synthetic workloads of the benchmark suite, built from seeds.

* :func:`network_2d`   one 2D unit-square network: N acceptors and M donors
  uniform from a seeded RandomState, the reference's 8 perimeter electrodes
  (thesis_indrek/generate_tests.py:32-40) with voltages U(-1, 1) kT/e, kT = nu =
  1, a = 0.1 (ab = a*R), I_0 = 50 kT (the reference's I_0*R/d convention,
  kmc_dopant_networks.py:66), eV_constant from the P1 Laplace solve, and N - M
  holes placed as place_charges_random does.
* :func:`ensemble`     the same network as K trajectories with Philox keys 0..K-1.
* :func:`iv_sweep`     electrode 0 swept over linspace(-2, 2, n_v), the other
  electrodes held, R replicas per voltage; energy sets by superposition of
  the unit-voltage Laplace basis.
* :func:`uneven_batch` seeded 3D layouts in a 1 x 1 x 0.2 box with N in
  [10, 120], M = round(N/10), 4-12 electrodes on the side faces at z = 0.1,
  log-uniform hop budgets (potential: the engine's slab model).
* :func:`request_stream` Poisson arrivals over 64 resident layouts.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import laplace
from . import tables as T

KT = 1.0
NU = 1.0


@dataclass
class Network:
    N: int
    P: int
    M: int
    nu: float
    kT: float
    I_0: float
    R: float
    ab: float
    acceptors: np.ndarray
    donors: np.ndarray
    electrodes: np.ndarray
    distances: np.ndarray
    transitions_constant: np.ndarray
    comp_constant: np.ndarray
    eV_constant: np.ndarray
    E_constant: np.ndarray
    site_energies: np.ndarray
    occupation: np.ndarray
    dims: tuple = (1.0, 1.0, 0.0)
    extra: dict = field(default_factory=dict)

    def oracle_args(self):
        return (self.N, self.P, self.nu, self.kT, self.I_0, self.R, self.distances,
                self.E_constant, self.site_energies, self.transitions_constant)


def perimeter_points(P: int, z: float = 0.0) -> np.ndarray:
    """P points evenly spaced along the unit square's perimeter (no corners)."""
    el = np.zeros((P, 4))
    for e in range(P):
        s = (e + 0.5) / P * 4.0
        side, f = int(s), s - int(s)
        el[e, :3] = [(f, 0, z), (1, f, z), (1 - f, 1, z), (0, 1 - f, z)][side]
    return el


def _finish(N, M, P, I_0, a, acc, don, el, dims, rng, eV) -> Network:
    xdim, ydim, zdim = dims
    _, R = T.characteristic_length(N, xdim, ydim, zdim)
    ab = a * R
    d = T.calc_distances(acc, el)
    tc = T.calc_transitions_constant(d, ab, NU)
    comp = T.calc_comp_constant(acc, don, I_0 * KT, R)
    E = eV + comp
    se = np.zeros(N + P)
    se[N:] = el[:, 3]
    occ = T.place_charges_random(rng, N, M)
    return Network(N, P, M, NU, KT, I_0 * KT, R, ab, acc, don, el, d, tc, comp, eV, E, se, occ,
                   dims)


def network_2d(seed: int = 0, N: int = 30, M: int = 3, I_0: float = 50.0, a: float = 0.1,
               vmax: float = 1.0, res: float = 0.01) -> Network:
    """The plain single network (FP64 CPU reference case) for a seed."""
    rng = np.random.RandomState(seed)
    acc = rng.rand(N, 3) * [1.0, 1.0, 0.0]
    don = rng.rand(M, 3) * [1.0, 1.0, 0.0]
    V = rng.uniform(-vmax, vmax, 8)
    el = T.perimeter_electrodes(1.0, 1.0, V)
    eV = laplace.eV_constant(acc, el, 1.0, 1.0, 0.0, res=res, mu=0.0)
    return _finish(N, M, 8, I_0, a, acc, don, el, (1.0, 1.0, 0.0), rng, eV)


@dataclass
class Ensemble:
    net: Network
    seeds: np.ndarray
    hops: int


def ensemble(seed: int = 0, K: int = 4096, hops: int = 100_000) -> Ensemble:
    """The 4096-trajectory ensemble on one layout, Philox keys 0..K-1."""
    return Ensemble(network_2d(seed), np.arange(K, dtype=np.uint64), hops)


@dataclass
class IVSweep:
    net: Network
    voltages: np.ndarray          # [n_v] electrode `electrode` voltages
    electrode: int
    replicas: int
    hops: int
    E_sets: np.ndarray            # [n_v, N]
    V_sets: np.ndarray            # [n_v, P]
    basis: laplace.LaplaceBasis


def iv_sweep(seed: int = 0, n_v: int = 1024, replicas: int = 64, hops: int = 200_000,
             electrode: int = 0, vrange=(-2.0, 2.0)) -> IVSweep:
    net = network_2d(seed)
    volts = np.linspace(vrange[0], vrange[1], n_v)
    basis = laplace.LaplaceBasis(net.acceptors, net.electrodes, 1.0, 1.0, res=0.01)
    Vs = np.repeat(net.electrodes[None, :, 3], n_v, axis=0)
    Vs[:, electrode] = volts
    E = basis.eV(Vs) + net.comp_constant[None, :]
    return IVSweep(net, volts, electrode, replicas, hops, E, Vs, basis)


@dataclass
class UnevenBatch:
    nets: list
    hops: np.ndarray
    seeds: np.ndarray


def uneven_batch(seed: int = 0, n: int = 2048, hop_range=(1e4, 1e6), N_range=(10, 120),
                 P_range=(4, 12), zdim: float = 0.2, I_0: float = 50.0, a: float = 0.1,
                 res: float = 0.01) -> UnevenBatch:
    rng = np.random.RandomState(seed)
    nets = []
    for _ in range(n):
        N = int(rng.randint(N_range[0], N_range[1] + 1))
        M = int(round(N / 10))
        P = int(rng.randint(P_range[0], P_range[1] + 1))
        acc = rng.rand(N, 3) * [1.0, 1.0, zdim]
        don = rng.rand(M, 3) * [1.0, 1.0, zdim]
        el = perimeter_points(P, zdim / 2)
        el[:, 3] = rng.uniform(-1, 1, P)
        eV = laplace.potential_3d_slab(acc, el, 1.0, 1.0, res=res)
        nets.append(_finish(N, M, P, I_0, a, acc, don, el, (1.0, 1.0, zdim), rng, eV))
    lo, hi = np.log(hop_range[0]), np.log(hop_range[1])
    hops = np.exp(rng.uniform(lo, hi, n)).astype(np.int64)
    return UnevenBatch(nets, hops, np.arange(n, dtype=np.uint64))


def network_3d(seed: int, N: int, P: int, zdim: float = 0.2, I_0: float = 50.0, a: float = 0.1,
               res: float = 0.01) -> Network:
    """One network built as the uneven batch builds its networks, with a given N
    and P (M = round(N / 10) donors)."""
    rng = np.random.RandomState(seed)
    M = int(round(N / 10))
    acc = rng.rand(N, 3) * [1.0, 1.0, zdim]
    don = rng.rand(M, 3) * [1.0, 1.0, zdim]
    el = perimeter_points(P, zdim / 2)
    el[:, 3] = rng.uniform(-1, 1, P)
    eV = laplace.potential_3d_slab(acc, el, 1.0, 1.0, res=res)
    return _finish(N, M, P, I_0, a, acc, don, el, (1.0, 1.0, zdim), rng, eV)


@dataclass
class Request:
    rid: int
    t_arrive: float
    layout: int
    V: np.ndarray
    hops: int


@dataclass
class RequestStream:
    layouts: list                 # Network objects (resident)
    bases: list                   # LaplaceBasis per layout
    requests: list


def resident_layouts(seed: int = 0, n: int = 64, N_range=(20, 100)) -> tuple[list, list]:
    rng = np.random.RandomState(seed)
    nets, bases = [], []
    for i in range(n):
        N = int(rng.randint(N_range[0], N_range[1] + 1))
        M = int(round(N / 10))
        net = network_2d(seed * 1000 + i, N=N, M=M)
        nets.append(net)
        bases.append(laplace.LaplaceBasis(net.acceptors, net.electrodes, 1.0, 1.0, res=0.01))
    return nets, bases


def request_stream(rate: float, n_req: int, layouts: list, seed: int = 1,
                   vmax: float = 1.5, median_hops: float = 5e4, sigma: float = 1.0,
                   max_hops: int = 2_000_000) -> list:
    rng = np.random.RandomState(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    lay = rng.randint(0, len(layouts), n_req)
    V = rng.uniform(-vmax, vmax, (n_req, 8))
    hops = np.minimum(np.exp(np.log(median_hops) + sigma * rng.randn(n_req)), max_hops)
    return [Request(i, float(t[i]), int(lay[i]), V[i], max(1, int(hops[i]))) for i in range(n_req)]


@dataclass
class IVRequest:
    rid: int
    t_arrive: float
    layout: int
    V: np.ndarray                 # (points, P): electrode `electrode` swept, the others held
    hops: int
    replicas: int


def iv_request_stream(rate: float, n_req: int, layouts: list, seed: int = 1, points: int = 16,
                      replicas: int = 4, hops: int = 200_000, electrode: int = 0,
                      vrange=(-2.0, 2.0), vmax: float = 1.5) -> list:
    """Poisson arrivals of IV-sweep requests (the evaluation of IV,
    kmc_dopant_networks_utils.py:639-693, as a served job): each names a resident
    layout, holds its other electrodes at U(-vmax, vmax) and sweeps `electrode`
    over linspace(vrange, points), `replicas` trajectories of `hops` hops per point."""
    rng = np.random.RandomState(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    lay = rng.randint(0, len(layouts), n_req)
    sweep = np.linspace(vrange[0], vrange[1], points)
    out = []
    for i in range(n_req):
        P = layouts[int(lay[i])].P
        V = np.repeat(rng.uniform(-vmax, vmax, P)[None], points, 0)
        V[:, electrode] = sweep
        out.append(IVRequest(i, float(t[i]), int(lay[i]), V, int(hops), int(replicas)))
    return out
