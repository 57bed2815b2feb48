"""Phase-timing instantiations of the FP64 dense and FP32 batched kernels.

They add shader-clock reads at the phase boundaries of every hop and change no
arithmetic: a batch run through kmc_probe_fp64_phases / kmc_probe_fp32_phases
ends in the same bits as the ordinary kernel's run, and the phase cycles are
positive with no more hits than hops.
"""
import ctypes as C

import numpy as np
import pytest

from kmc_dn_amd import _lib, engine, probe, workloads

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.time, b.time)
    assert np.array_equal(a.eo, b.eo)
    assert np.array_equal(a.occ_bits, b.occ_bits)
    assert np.array_equal(a.status, b.status)


def _net_batch(N, P, K, hops, seed=1):
    net = workloads.network_3d(seed, N, P)
    lay = engine.Layout(N, P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
    es = lay.energy_sets(net.E_constant, net.site_energies[N:])
    b = engine.Batch(engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), hops,
                                       net.occupation))
    return lay, b


def test_fp64_phase_instantiation_matches(gpu_engine):
    print("[gpu] FP64 phase timing, N=30 P=8, 4 trajectories x 3000 hops", flush=True)
    lay, b = _net_batch(30, 8, 4, 3000)
    try:
        b.run(precision=_lib.FP64_DENSE)
        b.sync()
        r0 = b.results()
        cyc = (C.c_uint64 * 8)()
        _lib.check(_lib.load().kmc_probe_fp64_phases(b.id, cyc, 8))
        _same(r0, b.results())
        assert all(cyc[k] > 0 for k in range(len(probe.FP64_PHASES)))
    finally:
        b.close()
        lay.close()


@pytest.mark.parametrize("N,P,form,team", [(30, 8, 1, 1), (30, 8, 2, 1), (90, 10, 1, 4),
                                           (120, 12, 2, 4), (150, 12, 3, 4)])
def test_fp32_phase_instantiation_matches(gpu_engine, N, P, form, team):
    K, hops = (256, 20_000) if team == 1 else (8, 5_000)
    print(f"[gpu] FP32 phase timing N={N} P={P} form {form} team {team}: {K} x {hops}", flush=True)
    lay, b = _net_batch(N, P, K, hops)
    try:
        run = dict(precision=_lib.FP32_BATCHED, table_form=form, team=team,
                   waves_per_group=4 if team == 1 else 1)
        b.run(**run)
        b.sync()
        r0 = b.results()
        o = _lib.kmc_run_opts(_lib.FP32_BATCHED, run["waves_per_group"], 0, 0, 0, 0, 0, form, 0,
                              team)
        cyc = (C.c_uint64 * 11)()
        _lib.check(_lib.load().kmc_probe_fp32_phases(b.id, C.byref(o), cyc, 11))
        _same(r0, b.results())
        total = K * hops
        assert 0 < cyc[7] <= total and 0 < cyc[8] <= total and cyc[9] <= cyc[8]
        for k in (0, 3, 4, 5):
            assert cyc[k] > 0, probe.FP32_PHASES[k]
        # fast hops and leaps: one-chunk kernels and table-form solo kernels
        if N + P <= 64 or (team == 1 and form != 3):
            assert cyc[6] > 0, probe.FP32_PHASES[6]
        else:
            assert cyc[6] == 0 and cyc[9] == 0
        if cyc[7] < total:
            assert cyc[1] > 0 and cyc[2] > 0
    finally:
        b.close()
        lay.close()


def test_fp32_phase_probe_rejects_shapes_it_has_no_kernel_for(gpu_engine):
    lay, b = _net_batch(30, 8, 4, 100)
    try:
        for form, w, team in [(3, 4, 1), (1, 2, 1)]:
            o = _lib.kmc_run_opts(_lib.FP32_BATCHED, w, 0, 0, 0, 0, 0, form, 0, team)
            cyc = (C.c_uint64 * 11)()
            assert _lib.load().kmc_probe_fp32_phases(b.id, C.byref(o), cyc, 11) != 0
        o = _lib.kmc_run_opts(_lib.FP32_BATCHED, 4, 0, 1, 0, 0, 0, 1, 0, 1)   # records
        assert _lib.load().kmc_probe_fp32_phases(b.id, C.byref(o), cyc, 11) != 0
        o = _lib.kmc_run_opts(_lib.FP32_BATCHED, 4, 0, 0, 0, 0, 0, 1, 0, 1)
        assert _lib.load().kmc_probe_fp32_phases(b.id, C.byref(o), cyc, 10) != 0   # slots
    finally:
        b.close()
        lay.close()
