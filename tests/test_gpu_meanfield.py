"""Mean-field estimators on the GPU against the oracle: probSimulate
(goSimulation/probabilitySimulation.go:53-157) and calcAverageCurrent
(goSimulation/simulation.go:671-736), through the C-ABI, the exact-name Go
exports (wrapperSimulateProbability, wrapperSimulateCombined) and the kmc_dn
class.  The kernel sums rows and columns instead of the reference's row-major
loops and uses the device exp: agreement to rounding (tolerances below)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from kmc_dn_amd import _lib, engine
from tests.netgen import fixture_network, random_network
from tests.test_gpu_goshim import GoSlice, go_slice, _read

pytestmark = pytest.mark.gpu


def _nets(fixtures):
    yield "fixture-rnd-0", fixture_network(fixtures, "rnd", 0)
    yield "N20-P6", random_network(1, 20, 6, I_0=50.0, a=0.1, vmax=1.0)
    yield "N8-P0", random_network(2, 8, 0, M=3, I_0=20.0, a=0.3)
    yield "N70-P10", random_network(3, 70, 10, I_0=50.0, a=0.1, vmax=1.5)
    yield "N120-P12", random_network(4, 120, 12, I_0=50.0, a=0.1, vmax=1.0)


def _layout(net):
    N, P = net["N"], net["P"]
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    return lay, es


def _oracle(net, steps, occ0, record):
    return oracle.prob_simulate(net["N"], net["P"], net["nu"], net["kT"], net["I_0"], net["R"],
                                net["distances"], net["E_constant"], net["site_energies"],
                                net["transitions_constant"], occ0, steps, record=record)


def test_meanfield_matches_oracle(fixtures, gpu_engine):
    steps = 200
    for name, net in _nets(fixtures):
        N, P, S = net["N"], net["P"], net["N"] + net["P"]
        print(f"[gpu] meanfield {name} S={S} steps={steps}", flush=True)
        lay, es = _layout(net)
        try:
            occ0 = np.full(N, 0.5)
            g = engine.meanfield_simulate([lay], es, steps, occ0, record=True)
            ref = _oracle(net, steps, occ0, True)
            print(f"[gpu]   time rel {abs(g.time[0] - ref.time) / ref.time:.1e}, occupation "
                  f"{np.abs(g.occupation[0, :N] - ref.occupation).max():.1e}", flush=True)
            assert g.time[0] == pytest.approx(ref.time, rel=1e-10)
            # occupations in [0, 1]; larger networks sum more terms per step
            assert np.allclose(g.occupation[0, :N], ref.occupation, rtol=0, atol=1e-8)
            assert np.allclose(g.electrode_occupation[0, :P], ref.electrode_occupation,
                               rtol=1e-9, atol=1e-8)
            assert np.allclose(g.site_energies[0, :N], ref.site_energies[:N], rtol=1e-9,
                               atol=1e-9)
            assert np.allclose(g.average_occupation[0, :N], ref.average_occupation, rtol=1e-9,
                               atol=1e-8 * ref.time)
            assert np.allclose(g.traffic[0, :S, :S], ref.traffic, rtol=1e-9, atol=1e-8)
        finally:
            lay.close()


def test_meanfield_batch_equals_single_networks(fixtures, gpu_engine):
    nets = [n for _, n in _nets(fixtures)][:4]
    lays, sets = zip(*[_layout(n) for n in nets])
    try:
        occ0 = np.zeros((len(nets), _lib.NMAX))
        for i, n in enumerate(nets):
            occ0[i, :n["N"]] = np.linspace(0.1, 0.9, n["N"])
        print("[gpu] meanfield batch of 4 layouts", flush=True)
        b = engine.meanfield_simulate(lays, [s[0] for s in sets], 150, occ0)
        for i, n in enumerate(nets):
            one = engine.meanfield_simulate([lays[i]], sets[i], 150, occ0[i:i + 1])
            assert np.array_equal(one.occupation[0], b.occupation[i])
            assert np.array_equal(one.electrode_occupation[0], b.electrode_occupation[i])
            assert one.time[0] == b.time[i]
    finally:
        for lay in lays:
            lay.close()


def test_average_current_matches_oracle(fixtures, gpu_engine):
    rng = np.random.RandomState(8)
    for name, net in _nets(fixtures):
        N, P = net["N"], net["P"]
        if P == 0:
            continue
        print(f"[gpu] average current {name}", flush=True)
        lay, es = _layout(net)
        try:
            occ = rng.uniform(0.05, 0.95, N)
            g = engine.average_current([lay], es, occ)[0, :P]
            ref = oracle.average_current(N, P, net["nu"], net["kT"], net["I_0"], net["R"],
                                         net["distances"], net["E_constant"],
                                         net["site_energies"], net["transitions_constant"], occ)
            scale = np.abs(ref).max()
            assert np.allclose(g, ref, rtol=1e-7, atol=1e-9 * scale)
        finally:
            lay.close()


def _go_fn(name):
    f = getattr(C.CDLL(_lib.LIB_PATH), name)
    # goSimulation/pythonBind.py:65-66: 7 doubles (time included), c_int hops
    f.argtypes = [C.c_longlong, C.c_longlong, C.c_double, C.c_double, C.c_double, C.c_double,
                  C.c_double, GoSlice, GoSlice, GoSlice, GoSlice, GoSlice, GoSlice, C.c_int,
                  C.c_bool, GoSlice, GoSlice]
    f.restype = C.c_double
    return f


def test_go_probability_and_combined_exports(gpu_engine):
    net = random_network(5, 20, 6, I_0=20.0, a=0.3, vmax=2.0)
    N, P = 20, 6
    S = N + P
    keep = []

    def gs(x):
        s, a = go_slice(np.ravel(x))
        keep.append(a)
        return s
    occ = gs(np.zeros(N))
    eo = gs(np.zeros(P))
    traffic = gs(np.zeros(S * S))
    avg = gs(np.zeros(N))
    se = gs(net["site_energies"])
    print("[gpu] wrapperSimulateProbability", flush=True)
    t = _go_fn("wrapperSimulateProbability")(
        N, P, 1.0, 1.0, net["I_0"], net["R"], 0.0, occ, gs(net["distances"]),
        gs(net["E_constant"]), gs(net["transitions_constant"]), eo, se, 300, True, traffic, avg)
    lay, es = _layout(net)
    try:
        ref = engine.meanfield_simulate([lay], es, 300, np.full(N, 0.5), record=True)
        assert t == ref.time[0]
        assert np.array_equal(_read(occ, N), ref.occupation[0, :N])    # written back, fractional
        assert np.array_equal(_read(eo, P), ref.electrode_occupation[0, :P])
        assert np.array_equal(_read(se, N), ref.site_energies[0, :N])
        assert np.array_equal(_read(traffic, S * S), ref.traffic[0].ravel())
        assert np.array_equal(_read(avg, N), ref.average_occupation[0, :N])
        print("[gpu] wrapperSimulateCombined", flush=True)
        eo2, tr2, avg2 = gs(np.zeros(P)), gs(np.zeros(S * S)), gs(np.zeros(N))
        t2 = _go_fn("wrapperSimulateCombined")(
            N, P, 1.0, 1.0, net["I_0"], net["R"], 0.0, gs(np.zeros(N)), gs(net["distances"]),
            gs(net["E_constant"]), gs(net["transitions_constant"]), eo2, gs(net["site_energies"]),
            20000, False, tr2, avg2)
        assert t2 > 0
        occ_time = _read(avg2, N)     # recorded whatever `record` says
        assert occ_time.sum() > 0 and (occ_time <= t2 * (1 + 1e-12)).all()
        cur = engine.average_current([lay], es, occ_time / t2)[0, :P]
        assert np.array_equal(_read(eo2, P), cur * t2)
        tr = _read(tr2, S * S).reshape(S, S)
        assert np.array_equal(tr, -tr.T) and np.abs(tr).sum() > 0
    finally:
        lay.close()


def test_kmc_class_estimators(gpu_engine):
    from kmc_dn_amd.kmc import kmc_dn
    el = np.zeros((4, 4))
    el[:, :2] = [[0.0, 0.25], [1.0, 0.75], [0.25, 0.0], [0.75, 1.0]]
    el[:, 3] = [1.0, -1.0, 0.5, 0.0]
    k = kmc_dn(12, 2, 1.0, 1.0, 0, electrodes=el, I_0=20.0, a=0.3, res=0.05)
    print("[gpu] kmc_dn.go_simulation(wrapperSimulateProbability)", flush=True)
    k.go_simulation(hops=300, goSpecificFunction="wrapperSimulateProbability", record=True)
    est = k.estimator
    assert np.array_equal(k.electrode_occupation, np.array([int(x) for x in est.electrode_occupation]))
    assert np.array_equal(k.current, k.electrode_occupation / k.time)
    assert len(k.average_occupation) == 12 and k.traffic.shape == (16, 16)
    print("[gpu] kmc_dn.go_simulation(wrapperSimulateCombined)", flush=True)
    k.go_simulation(hops=5000, goSpecificFunction="wrapperSimulateCombined")
    assert k.time > 0 and np.isfinite(k.estimator.electrode_occupation).all()
