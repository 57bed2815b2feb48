"""Device time limits of the FP32 batched kernel (kmc_run_opts.time_limit_us).

A trajectory that has run past the limit stops at its next recompute point, a
multiple of the recompute interval, with status OK and hops_done short of its
range; resumed from there it continues exactly as one uninterrupted launch
would.  The serving loop ends its rounds this way.  No reference counterpart:
the reference runs each call to its end (kmc_dopant_networks.py:33-135).
"""
import numpy as np
import pytest

from kmc_dn_amd import _lib, engine, workloads

pytestmark = pytest.mark.gpu


def _ensemble(K, hops, seed=0):
    ens = workloads.ensemble(seed, K, hops)
    net = ens.net
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
    es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])
    return lay, engine.make_trajs(K, lay.id, es[0], ens.seeds, hops, net.occupation)


def _bits(r):
    return (r.time.view(np.uint64).copy(), r.eo.copy(), r.occ_bits.copy(), r.status.copy())


@pytest.mark.parametrize("form", [1, 3])
def test_stopped_launches_resume_to_the_uninterrupted_result(gpu_engine, form):
    K, hops, limit_us = 256, 200_000, 500
    print(f"[gpu] time limit {limit_us} us, form {form}: {K} x {hops}", flush=True)
    lay, trajs = _ensemble(K, hops)
    try:
        b = engine.Batch(trajs)
        b.run(precision=_lib.FP32_BATCHED, table_form=form)
        b.sync()
        whole = _bits(b.results())
        b.close()

        b = engine.Batch(trajs)
        reached = np.zeros(K, np.int64)
        launches = 0
        stopped_early = 0
        while (reached < hops).any():
            b.set_state(h_begin=reached, h_end=np.full(K, hops, np.int64))
            b.run(precision=_lib.FP32_BATCHED, table_form=form, resume=True,
                  time_limit_us=limit_us)
            b.sync()
            r = b.results()
            assert (r.status == 0).all()
            assert (r.hops_done >= reached).all()
            short = r.hops_done < hops
            # a stop falls on a recompute point (the default interval, 1024 hops)
            assert (r.hops_done[short] % 1024 == 0).all()
            stopped_early += int(short.sum())
            reached = r.hops_done.copy()
            launches += 1
            assert launches < 500
        print(f"[gpu] {launches} launches, {stopped_early} early stops", flush=True)
        assert launches > 1 and stopped_early > 0
        pieces = _bits(b.results())
        b.close()
        for x, y in zip(whole, pieces):
            assert np.array_equal(x, y)
    finally:
        lay.close()


def test_time_limits_are_refused_where_unsupported(gpu_engine):
    lay, trajs = _ensemble(8, 1000)
    try:
        b = engine.Batch(trajs)
        with pytest.raises(_lib.KmcError):
            b.run(precision=_lib.FP32_BATCHED, team=4, time_limit_us=100)
        with pytest.raises(_lib.KmcError):
            b.run(precision=_lib.FP64_DENSE, time_limit_us=100)
        with pytest.raises(_lib.KmcError):
            b.run(precision=_lib.FP32_BATCHED, time_limit_us=-1)
        b.run(precision=_lib.FP32_BATCHED, time_limit_us=100)   # solo: accepted
        b.sync()
        assert (b.results().status == 0).all()
        b.close()
    finally:
        lay.close()


def test_watchdog_wait(gpu_engine):
    """kmc_batch_wait returns KMC_ERR_TIMEOUT (None here) while a launch runs and
    leaves it running; a later wait collects it with the kernel time of sync."""
    lay, trajs = _ensemble(4096, 300_000)
    try:
        b = engine.Batch(trajs)
        b.run(precision=_lib.FP32_BATCHED)
        first = b.wait(0.0)
        ms = b.wait(60_000.0)
        print(f"[gpu] watchdog: first wait {first}, kernel {ms:.1f} ms", flush=True)
        assert ms is not None and ms > 0
        if ms > 20.0:        # a launch that long cannot have ended at once
            assert first is None
        assert b.sync() == ms
        assert (b.results().status == 0).all()
        with pytest.raises(_lib.KmcError):
            _lib.check(_lib.load().kmc_batch_wait(b.id, -1.0, None))
        b.close()
    finally:
        lay.close()
