"""Uneven batches: size buckets, autotuned plans, graph capture, statistics."""
import os

import numpy as np
import pytest

from tests.stats import assert_within, compare

from kmc_dn_amd import _lib, autotune, engine, uneven, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_uneven(gpu_engine):
    ub = workloads.uneven_batch(seed=3, n=192, hop_range=(1e3, 2e4))
    dev = uneven.upload(ub)
    yield ub, dev
    dev.close()


def _same(a, b):
    assert np.array_equal(a.eo, b.eo)
    assert np.array_equal(a.time, b.time)
    assert np.array_equal(a.occ_bits, b.occ_bits)


def test_bucketed_positions_equal_one_kernel(small_uneven):
    """The positions form gives the same bits whatever the site-chunk count of the
    kernel that runs a network (chunks past S are skipped), so one C=3 kernel, a
    three-bucket plan and different workgroup sizes agree exactly."""
    ub, dev = small_uneven
    S = np.array([n.N + n.P for n in ub.nets])
    assert S.min() <= 64 < S.max()
    print("[gpu] uneven: one positions kernel", flush=True)
    one = uneven.run(dev, plan=[(192, 3, 4, 0)], use_graph=False)
    uneven.check_invariants(dev, one.result)
    for plan in ([(64, 3, 1, 0), (128, 3, 2, 0), (192, 3, 8, 0)],
                 [(128, 3, 4, 0), (192, 3, 1, 0)]):
        print(f"[gpu] uneven: plan {plan}", flush=True)
        r = uneven.run(dev, plan=plan, use_graph=False)
        _same(r.result, one.result)


def test_graph_replays_match_eager(small_uneven):
    ub, dev = small_uneven
    plan = [(64, 1, 4, 0), (128, 2, 2, 0), (192, 3, 4, 0)]
    print("[gpu] uneven: eager vs graph replays", flush=True)
    eager = uneven.run(dev, plan=plan, use_graph=False)
    uneven.check_invariants(dev, eager.result)
    b = engine.Batch(dev.trajs)
    try:
        b.set_buckets(plan)
        for rep in range(3):
            b.run(precision=_lib.FP32_BATCHED, use_graph=True)
            b.sync()
            _same(b.results(), eager.result)
            # eager work between the replays
            b.run(precision=_lib.FP32_BATCHED, use_graph=False)
            b.sync()
    finally:
        b.close()


def test_mixed_forms_statistics(gpu_engine):
    """Table forms and the positions form evaluate the same rates with different
    FP32 roundings: ensemble currents agree within their standard errors."""
    ub = workloads.uneven_batch(seed=5, n=12, hop_range=(2e4, 2e4 + 1), N_range=(10, 110))
    reps = 48
    nets = ub.nets
    lays, trajs, sizes = [], [], {}
    for i, net in enumerate(nets):
        lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                            net.transitions_constant)
        lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
        es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])[0]
        trajs.append(engine.make_trajs(reps, lay.id, es, np.arange(reps, dtype=np.uint64) + 1000 * i,
                                       20_000, net.occupation))
        lays.append(lay)
        sizes[lay.id] = (lay.S, lay.N)
    t = np.concatenate(trajs)
    out = {}
    try:
        for name, plan in [("pos", [(192, 3, 4, 0)]),
                           ("tab", [(64, 1, 4, 0), (128, 2, 2, 0)])]:
            print(f"[gpu] mixed forms: {name}", flush=True)
            b = engine.Batch(t)
            b.set_buckets(plan)
            b.run(precision=_lib.FP32_BATCHED)
            b.sync()
            r = b.results()
            b.close()
            assert (r.status == 0).all()
            out[name] = r
    finally:
        for lay in lays:
            lay.close()
    cs = [compare(out["pos"].currents(net.P)[i * reps:(i + 1) * reps],
                  out["tab"].currents(net.P)[i * reps:(i + 1) * reps]) for i, net in enumerate(nets)]
    assert_within({k: np.concatenate([c[k] for c in cs]) for k in ("z", "resolvable", "rel")},
                  f"positions vs tables over {len(nets)} networks")


def test_uneven_fp32_vs_fp64(gpu_engine):
    """FP32 positions form against the FP64 dense path on networks with S <= 80."""
    ub = workloads.uneven_batch(seed=7, n=6, hop_range=(3e4, 3e4 + 1), N_range=(10, 60))
    reps = 64
    zs = []
    for i, net in enumerate(ub.nets):
        lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                            net.transitions_constant)
        lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
        es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])[0]
        seeds = np.arange(reps, dtype=np.uint64) + 7000 * i
        cur = {}
        for prec, form in [(_lib.FP64_DENSE, 0), (_lib.FP32_BATCHED, 3)]:
            print(f"[gpu] uneven net {i} N={net.N} P={net.P} precision {prec}", flush=True)
            b = engine.Batch(engine.make_trajs(reps, lay.id, es, seeds, 30_000, net.occupation))
            b.run(precision=prec, table_form=form)
            b.sync()
            r = b.results()
            b.close()
            assert (r.status == 0).all()
            cur[prec] = r.currents(net.P)
        lay.close()
        zs.append(compare(cur[_lib.FP32_BATCHED], cur[_lib.FP64_DENSE]))
    assert_within({k: np.concatenate([c[k] for c in zs]) for k in ("z", "resolvable", "rel")},
                  "fp32 positions vs fp64 over 6 uneven networks")


def test_autotune_plan_and_cache(small_uneven, tmp_path):
    ub, dev = small_uneven
    path = str(tmp_path / "tune.json")
    print("[gpu] autotune", flush=True)
    plan = autotune.tune(dev.trajs, dev.sizes, sample=128, sample_hops=1024, cache_path=path)
    S = np.array([dev.sizes[int(x)][0] for x in dev.trajs["layout"]])
    edges = [p[0] for p in plan]
    assert edges == sorted(edges) and edges[-1] >= S.max() and len(edges) >= 2
    assert all(p[1] in (1, 2, 3) and p[2] in (1, 2, 4, 8) for p in plan)
    assert os.path.exists(path)
    again = autotune.tune(dev.trajs, dev.sizes, cache_path=path)   # read back, no timing
    assert again == plan
    r = uneven.run(dev, plan=plan, use_graph=True)
    uneven.check_invariants(dev, r.result)
