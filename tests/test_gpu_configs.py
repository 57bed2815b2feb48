"""The benchmark configurations themselves, built by kmc_dn_amd.workloads and
checked against the oracle.

* C0, the plain N=30 network with its P1 Laplace potential: the FP64 kernel
  replays 1e4 hops of a uniform stream event for event against the numba-loop
  restatement (kmc_dopant_networks.py:33-135), then runs 1e5 hops of
  simulate_discrete on the Philox stream, again event for event.
* C1, the 4096-trajectory ensemble at 1e5 hops: the FP64 kernel's first 16
  trajectories equal the oracle's ensemble driver on the same Philox keys, and
  the FP32 kernel's per-electrode ensemble means lie within 3 sigma of the
  FP64 kernel's on all 4096.
* C2, the IV sweep: 16 of its 1024 voltages x 64 replicas x 2e5 hops, FP32
  against FP64 within 3 sigma at every (voltage, electrode).
"""
import numpy as np
import pytest

import oracle
from kmc_dn_amd import _lib, engine, iv, workloads
from tests.stats import assert_within, compare

pytestmark = pytest.mark.gpu


def _layout(net):
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])[0]
    return lay, es


def test_c0_replay_and_philox_runs_are_bit_exact(gpu_engine):
    net = workloads.network_2d(0)
    N, P = net.N, net.P
    for hops, uniforms in [(10_000, np.random.RandomState(5).rand(20_000)), (100_000, None)]:
        print(f"[gpu] C0 FP64 {hops} hops, {'replayed' if uniforms is not None else 'Philox'}",
              flush=True)
        ref = oracle.numba_simulate(*net.oracle_args(), net.occupation, np.zeros(P, np.int64),
                                    hops, uniforms=uniforms, seed=0, log=True, record=True)
        se = net.site_energies.copy()
        occ = net.occupation.copy()
        eo = np.zeros(P, np.int64)
        t, traffic, oit, ev, tot = engine.simulate_discrete(
            N, P, net.nu, net.kT, net.I_0, net.R, net.distances, net.E_constant, se,
            net.transitions_constant, occ, eo, hops, record=True, uniforms=uniforms, seed=0,
            log=True)
        assert ref.status == 0
        assert np.array_equal(ev, ref.events)
        assert np.array_equal(eo, ref.electrode_occupation)
        assert np.array_equal(occ, ref.occupation)
        assert np.array_equal(traffic, ref.traffic)
        assert abs(t - ref.time) <= 1e-12 * ref.time
        assert np.allclose(tot, ref.totals, rtol=1e-13, atol=0)
        assert np.allclose(oit, ref.occupations_in_time, rtol=1e-11, atol=0)
        assert np.allclose(se[:N], ref.site_energies[:N], rtol=1e-13, atol=1e-13)


def test_c1_ensemble_against_fp64_and_oracle(gpu_engine):
    ens = workloads.ensemble(0, 4096, 100_000)
    net = ens.net
    lay, es = _layout(net)
    res = {}
    for prec in (_lib.FP32_BATCHED, _lib.FP64_DENSE):
        print(f"[gpu] C1 4096 x 1e5 hops, precision {prec}", flush=True)
        b = engine.Batch(engine.make_trajs(4096, lay.id, es, ens.seeds, ens.hops,
                                           net.occupation))
        b.run(precision=prec)
        b.sync()
        res[prec] = b.results()
        b.close()
        assert (res[prec].status == 0).all() and (res[prec].hops_done == ens.hops).all()
    # the FP64 kernel is the oracle's arithmetic: the same trajectories, hop for hop
    k = 16
    eo, tm, occ, st = oracle.numba_ensemble(*net.oracle_args(), net.occupation, ens.seeds[:k],
                                            ens.hops)
    r64 = res[_lib.FP64_DENSE]
    assert (st == 0).all()
    assert np.array_equal(r64.eo[:k, :net.P], eo)
    assert np.array_equal(engine.bits_to_occupation(r64.occ_bits[:k], net.N), occ.astype(bool))
    assert np.allclose(r64.time[:k], tm, rtol=1e-12, atol=0)
    c = compare(res[_lib.FP32_BATCHED].currents(net.P), r64.currents(net.P))
    print(f"[gpu] C1: electrode events per trajectory "
          f"{np.abs(r64.eo[:, :net.P]).sum(1).mean():.2f}", flush=True)
    assert_within(c, "C1 FP32 vs FP64 (4096 x 1e5)")
    lay.close()


def test_c2_iv_sweep_subset(gpu_engine):
    sw = workloads.iv_sweep(0, 1024, 64, 200_000)
    net = sw.net
    pick = np.linspace(0, 1023, 16).astype(int)
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    means = {}
    for prec in (_lib.FP32_BATCHED, _lib.FP64_DENSE):
        print(f"[gpu] C2 16 points x 64 replicas x 2e5 hops, precision {prec}", flush=True)
        b, ids = iv.iv_batch(lay, sw.E_sets[pick], sw.V_sets[pick], sw.replicas, sw.hops,
                             net.occupation, seed0=0)
        b.run(precision=prec)
        b.sync()
        r = b.results()
        b.close()
        engine.Layout.drop_energy_sets(ids)
        assert (r.status == 0).all()
        means[prec] = r.currents(net.P).reshape(len(pick), sw.replicas, net.P)
    zs = [compare(means[_lib.FP32_BATCHED][i], means[_lib.FP64_DENSE][i])
          for i in range(len(pick))]
    c = {k: np.stack([z[k] for z in zs]) for k in zs[0]}
    assert_within(c, "C2 FP32 vs FP64 (16 of 1024 voltages)")
    lay.close()
