"""FP32-rate / FP64-accumulate batched kernel: determinism, invariants, statistics."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, engine
from tests.netgen import fixture_network, random_network
from tests.stats import assert_within, compare

pytestmark = pytest.mark.gpu


def _setup(net, K, hops, seed0=0, prehops=0):
    N, P = net["N"], net["P"]
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    seeds = np.arange(seed0, seed0 + K, dtype=np.uint64)
    return lay, engine.Batch(engine.make_trajs(K, lay.id, es[0], seeds, hops, net["occupation"],
                                               prehops=prehops))


def test_results_independent_of_grouping_and_graph(gpu_engine):
    net = random_network(0, 30, 8, M=3, I_0=50.0, a=0.1, vmax=1.0)
    _, b = _setup(net, 100, 2000)
    outs = []
    for W, graph in [(1, False), (2, False), (4, False), (8, False), (4, True), (4, True)]:
        print(f"[gpu] fp32 W={W} graph={graph}", flush=True)
        b.run(precision=_lib.FP32_BATCHED, waves_per_group=W, use_graph=graph)
        b.sync()
        outs.append(b.results())
    for r in outs[1:]:
        assert np.array_equal(r.eo, outs[0].eo)
        assert np.array_equal(r.time, outs[0].time)
        assert np.array_equal(r.occ_bits, outs[0].occ_bits)


def test_invariants(gpu_engine):
    net = random_network(2, 40, 6, I_0=50.0, a=0.1, vmax=1.5)
    N, P = 40, 6
    _, b = _setup(net, 256, 5000)
    b.run(precision=_lib.FP32_BATCHED)
    b.sync()
    r = b.results()
    assert (r.status == 0).all() and (r.hops_done == 5000).all()
    holes0 = net["occupation"].sum()
    holes = engine.bits_to_occupation(r.occ_bits, N).sum(axis=1)
    # every hop moves one hole: what the electrodes absorbed left the acceptors
    assert np.array_equal(r.eo[:, :P].sum(axis=1), holes0 - holes)
    assert (r.eo[:, P:] == 0).all()
    assert (r.time > 0).all()


def _ensemble(net, precision, K, hops, **kw):
    _, b = _setup(net, K, hops, prehops=kw.pop("prehops", 0))
    b.run(precision=precision, **kw)
    ms = b.sync()
    r = b.results()
    assert (r.status == 0).all()
    return r.currents(net["P"]), ms


def _compare(ca, cb, label):
    """Every per-electrode ensemble mean within 3 sigma (tests/stats.py)."""
    assert_within(compare(ca, cb), label)


def test_fp32_matches_fp64_statistics_c1_like(gpu_engine):
    net = random_network(0, 30, 8, M=3, I_0=50.0, a=0.1, vmax=1.0)
    c64, ms64 = _ensemble(net, _lib.FP64_DENSE, 512, 20000, prehops=2000)
    c32, ms32 = _ensemble(net, _lib.FP32_BATCHED, 2048, 20000, prehops=2000)
    print(f"[gpu] fp64 {512 * 22000 / ms64 * 1e3:.3e} hops/s, fp32 {2048 * 22000 / ms32 * 1e3:.3e}"
          " hops/s", flush=True)
    _compare(c32, c64, "fp32 vs fp64 (C1-like)")


def test_recompute_interval_does_not_bias(gpu_engine):
    net = random_network(4, 50, 8, I_0=50.0, a=0.1, vmax=1.0)
    c1, _ = _ensemble(net, _lib.FP32_BATCHED, 1024, 10000, recompute_interval=1)
    c2, _ = _ensemble(net, _lib.FP32_BATCHED, 1024, 10000, recompute_interval=1 << 20)
    _compare(c1, c2, "recompute every hop vs never")


@pytest.mark.parametrize("N,P", [(60, 10), (100, 12), (120, 12)])
def test_multi_chunk_networks(gpu_engine, N, P):
    net = random_network(N, N, P, I_0=50.0, a=0.1, vmax=1.0)
    c32, _ = _ensemble(net, _lib.FP32_BATCHED, 512, 5000, prehops=1000)
    if N + P <= 80:
        c64, _ = _ensemble(net, _lib.FP64_DENSE, 256, 5000, prehops=1000)
        _compare(c32, c64, f"fp32 vs fp64 N={N} P={P}")
    else:
        assert np.isfinite(c32).all()


def test_fixture_currents_against_reference(fixtures, gpu_engine):
    """Ensemble means vs the reference's own mean/std over 5 runs of 1e6 hops."""
    zs = []
    for lset, sname in [("rnd", "rnd_min_max"), ("xor", "XOR_wide")]:
        for k in range(0, 100, 10):
            net = fixture_network(fixtures, lset, k)
            c, _ = _ensemble(net, _lib.FP32_BATCHED, 64, 100000)
            m, se = c.mean(0), c.std(0, ddof=1) / np.sqrt(len(c))
            rm = fixtures[f"{sname}_mean_currents"][k]
            rs = fixtures[f"{sname}_stddev_currents"][k]
            z = np.abs(m - rm) / np.sqrt(se ** 2 + rs ** 2 / 4)
            zs.append(z[np.abs(rm) > 1e-5])
    z = np.concatenate(zs)
    print(f"[gpu] fixtures: median |z| {np.median(z):.2f}, 95% {np.percentile(z, 95):.2f}, "
          f"frac>3 {np.mean(z > 3):.3f}", flush=True)
    # the reference's own numba-vs-Go comparison: median 0.74, p95 2.6, 3% above 3
    assert np.median(z) < 1.2 and np.mean(z > 3) < 0.08


def test_traffic_is_net_and_matches_tallies(gpu_engine):
    net = random_network(8, 24, 6, I_0=20.0, a=0.3, vmax=2.0)
    N, P = 24, 6
    _, b = _setup(net, 32, 4000)
    b.run(precision=_lib.FP32_BATCHED, traffic=True)
    b.sync()
    r = b.results()
    for j, tr in enumerate(b.traffic(N + P)):
        assert np.array_equal(tr, -tr.T)                       # net, antisymmetric
        into_el = tr[:, N:].sum(axis=0)                          # net hops into each electrode
        assert np.array_equal(into_el, r.eo[j, :P].astype(float))
        assert np.abs(tr).sum() > 0


def test_table_forms_agree(gpu_engine):
    net = random_network(12, 40, 8, I_0=30.0, a=0.2, vmax=2.0)
    ca, _ = _ensemble(net, _lib.FP32_BATCHED, 1024, 10000, table_form=1)
    cb, _ = _ensemble(net, _lib.FP32_BATCHED, 1024, 10000, table_form=2)
    _compare(ca, cb, "table {lt,lt+c} vs {lt,c}")


def _positions(net):
    from kmc_dn_amd import tables as T
    return T.site_positions(net["acceptors"], net["electrodes"])


def test_positions_form_matches_table_form(gpu_engine):
    net = random_network(21, 50, 8, I_0=30.0, a=0.25, vmax=2.0)
    N, P = 50, 8
    lay = engine.Layout(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    lay.set_positions(0.25 * net["R"], _positions(net))
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    b = engine.Batch(engine.make_trajs(1024, lay.id, es[0], np.arange(1024), 10000,
                                       net["occupation"], prehops=1000))
    out = {}
    for form in (1, 3):
        b.run(precision=_lib.FP32_BATCHED, table_form=form)
        b.sync()
        r = b.results()
        assert (r.status == 0).all()
        out[form] = r.currents(P)
    _compare(out[3], out[1], "positions vs {lt,lt+c} table")


def test_resume_in_chunks_is_bit_exact(gpu_engine):
    net = random_network(22, 30, 8, I_0=30.0, a=0.25, vmax=2.0)
    N, P = 30, 8
    K, hops, pre = 64, 9000, 500
    for form in (1, 3):
        lay = engine.Layout(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                            net["transitions_constant"])
        lay.set_positions(0.25 * net["R"], _positions(net))
        es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
        t = engine.make_trajs(K, lay.id, es[0], np.arange(K), hops, net["occupation"], prehops=pre)
        whole = engine.Batch(t)
        whole.run(precision=_lib.FP32_BATCHED, table_form=form, record=True)
        whole.sync()
        ref = whole.results()
        b = engine.Batch(t)
        total = hops + pre
        done = np.zeros(K, np.int64)
        chunk = 2048                           # a multiple of the recompute interval
        first = True
        while (done < total).any():
            end = np.minimum(done + chunk, total)
            b.set_state(h_begin=done, h_end=end)
            b.run(precision=_lib.FP32_BATCHED, table_form=form, record=True, resume=True)
            b.sync()
            done = end
            first = False
        r = b.results()
        assert not first
        assert np.array_equal(r.eo, ref.eo) and np.array_equal(r.time, ref.time), form
        assert np.array_equal(r.occ_bits, ref.occ_bits)
        assert np.array_equal(r.occ_time, ref.occ_time)
        assert (r.hops_done == hops).all()


def _team_checks(net, N, P, forms, K, hops):
    """Teams of 2 or 4 waves per trajectory: repeated and graph runs agree bit for
    bit, a chunked resume run equals the unchunked run, and the currents agree with
    the solo positions kernel statistically."""
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    lay.set_positions(0.1 * net["R"], _positions(net))
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    trajs = engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), hops,
                              net["occupation"])
    b = engine.Batch(trajs)
    cur = {}
    for form in forms:
        for team in (1, 2, 4):
            print(f"[gpu] S={N + P} team={team} form={form}", flush=True)
            b.run(precision=_lib.FP32_BATCHED, table_form=form, team=team)
            b.sync()
            r = b.results()
            assert (r.status == 0).all() and (r.hops_done == hops).all()
            if team > 1:
                b.run(precision=_lib.FP32_BATCHED, table_form=form, team=team, use_graph=True)
                b.sync()
                g = b.results()
                assert np.array_equal(g.eo, r.eo) and np.array_equal(g.time, r.time)
                # the same trajectories in three resumed chunks
                for lo, hi in ((0, 2048), (2048, 4096), (4096, hops)):
                    b.set_state(h_begin=np.full(K, lo), h_end=np.full(K, hi))
                    b.run(precision=_lib.FP32_BATCHED, table_form=form, team=team, resume=True)
                    b.sync()
                c = b.results()
                assert np.array_equal(c.eo, r.eo) and np.array_equal(c.time, r.time)
                assert np.array_equal(c.occ_bits, r.occ_bits)
                b.set_state(h_begin=np.zeros(K), h_end=np.full(K, hops))
            cur[(form, team)] = r.currents(P)
    b.close()
    lay.close()
    base = cur[(3, 1)]
    for key, c in cur.items():
        if key != (3, 1):
            _compare(c, base, f"form/team {key} vs positions solo")


def test_team_kernels_deterministic_and_resumable(gpu_engine):
    """Two site chunks (S = 80): the teams add the row sums in another order."""
    net = random_network(4, 70, 10, I_0=50.0, a=0.1, vmax=1.5)
    _team_checks(net, 70, 10, (1, 2, 3), K=96, hops=6144)


def test_three_chunk_team_kernels(gpu_engine):
    """Three site chunks (S = 134, whose FP32 table still fits LDS): the team
    reduction of three row-sum chunks per lane."""
    net = random_network(14, 122, 12, I_0=50.0, a=0.1, vmax=1.5)
    _team_checks(net, 122, 12, (2, 3), K=96, hops=4096)


def test_bucket_size_lower_bound(gpu_engine):
    """s_min bounds a bucket from below: the larger layout's trajectories take the
    team entry in front, the smaller one's fall through to the solo entry, each
    bit for bit as in a batch of its own; s_min > s_max is refused."""
    lays, parts = [], []
    for seed, N, P in ((21, 40, 8), (22, 80, 8)):   # S = 48 and 88
        net = random_network(seed, N, P, I_0=50.0, a=0.1, vmax=1.0)
        lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                            net["transitions_constant"])
        es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
        lays.append(lay)
        parts.append(engine.make_trajs(32, lay.id, es[0], np.arange(32, dtype=np.uint64) + 100 * seed,
                                       3000, net["occupation"]))
    b = engine.Batch(np.concatenate(parts))
    print("[gpu] s_min: S >= 60 -> team 4, the rest solo", flush=True)
    b.set_buckets([(128, 2, 1, 4, 0, 0, 60), (128, 2, 4, 1, 0)])
    b.run(precision=_lib.FP32_BATCHED, use_graph=True)
    b.sync()
    mixed = b.results()
    with pytest.raises(_lib.KmcError) as e:
        b.set_buckets([(64, 2, 1, 4, 0, 0, 65)])
    assert e.value.code == _lib.ERR_ARGS
    b.close()
    for i, team in ((0, 1), (1, 4)):
        one = engine.Batch(parts[i])
        one.run(precision=_lib.FP32_BATCHED, table_form=2, team=team)
        one.sync()
        r = one.results()
        one.close()
        sel = slice(32 * i, 32 * i + 32)
        assert (r.status == 0).all()
        assert np.array_equal(r.eo, mixed.eo[sel]) and np.array_equal(r.time, mixed.time[sel])
        assert np.array_equal(r.occ_bits, mixed.occ_bits[sel])
    for lay in lays:
        lay.close()


def test_team_kernel_chunk_count_is_irrelevant(gpu_engine):
    """An S = 100 layout's trajectories in a team entry whose largest network has
    S = 130 (a kernel of three site chunks) equal, bit for bit, the same
    trajectories in a batch of their own (two site chunks): the third chunk's
    guards add nothing (the team entries of a plan cut by LDS run so)."""
    lays, parts = [], []
    for seed, N, P in ((31, 90, 10), (32, 118, 12)):   # S = 100 and 130
        net = random_network(seed, N, P, I_0=50.0, a=0.1, vmax=1.0)
        lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                            net["transitions_constant"])
        es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
        lays.append(lay)
        parts.append(engine.make_trajs(16, lay.id, es[0], np.arange(16, dtype=np.uint64) + 100 * seed,
                                       3000, net["occupation"]))
    print("[gpu] S=100 trajectories in a three-chunk team kernel", flush=True)
    b = engine.Batch(np.concatenate(parts))
    b.set_buckets([(192, 2, 1, 4, 0)])
    b.run(precision=_lib.FP32_BATCHED, use_graph=True)
    b.sync()
    mixed = b.results()
    b.close()
    one = engine.Batch(parts[0])
    one.run(precision=_lib.FP32_BATCHED, table_form=2, team=4)
    one.sync()
    r = one.results()
    one.close()
    assert (r.status == 0).all() and (mixed.status == 0).all()
    assert np.array_equal(r.eo, mixed.eo[:16]) and np.array_equal(r.time, mixed.time[:16])
    assert np.array_equal(r.occ_bits, mixed.occ_bits[:16])
    for lay in lays:
        lay.close()


def test_budget_buckets_route_long_trajectories_to_teams(gpu_engine):
    net = random_network(6, 40, 8, I_0=50.0, a=0.1, vmax=1.0)
    N, P = 40, 8
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    K = 64
    hops = np.where(np.arange(K) % 4 == 0, 8192, 2048)
    trajs = engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), hops,
                              net["occupation"])
    b = engine.Batch(trajs)
    print("[gpu] budget buckets: long -> team 4, short -> solo", flush=True)
    b.set_buckets([(64, 1, 1, 4, 8192), (64, 1, 4, 1, 0)])
    b.run(precision=_lib.FP32_BATCHED, use_graph=True)
    b.sync()
    mixed = b.results()
    b.set_buckets([])
    for team, sel in ((4, hops == 8192), (1, hops == 2048)):
        b.run(precision=_lib.FP32_BATCHED, table_form=1, team=team)
        b.sync()
        r = b.results()
        assert np.array_equal(r.eo[sel], mixed.eo[sel]) and np.array_equal(r.time[sel], mixed.time[sel])
    b.close()
    lay.close()


def test_pruned_pairs_never_hop(gpu_engine):
    """wrapperSimulatePruned semantics (simulation.go:199-215): pairs with
    tc <= prune * max(tc) are dropped from the FP32 pair set."""
    net = random_network(21, 30, 8, I_0=50.0, a=0.25, vmax=1.0)
    N, P, S = 30, 8, 38
    tc = net["transitions_constant"]
    prune = 1e-3
    pruned = tc <= prune * tc.max()
    np.fill_diagonal(pruned, False)
    assert 0.05 < pruned.mean() < 0.95
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"], tc,
                        prune_threshold=prune)
    with pytest.raises(_lib.KmcError) as e:
        lay.set_positions(0.25 * net["R"], _positions(net))
    assert e.value.code == _lib.ERR_UNSUPPORTED
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    K = 64
    b = engine.Batch(engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), 20_000,
                                       net["occupation"]))
    for form in (1, 2):
        print(f"[gpu] pruned pair set, table form {form}", flush=True)
        b.run(precision=_lib.FP32_BATCHED, table_form=form, traffic=True)
        b.sync()
        assert (b.results().status == 0).all()
        tr = np.stack(b.traffic(S))          # net traffic, antisymmetric
        moved = np.abs(tr).sum(0) > 0
        assert moved.any() and not (moved & pruned).any()
    b.close()
    lay.close()


def test_electrode_closed_form_and_member_fallback(gpu_engine):
    """Table forms add the electrode rates in closed form while every scaled
    voltage |V log2(e) / kT| <= 100 and walk the electrodes as members otherwise;
    the positions form always walks them.  One batch mixes both cases: each
    agrees with the positions form statistically, and the mix changes no bits."""
    net = random_network(23, 40, 8, I_0=30.0, a=0.2, vmax=2.0)
    N, P = 40, 8
    lay = engine.Layout(N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    lay.set_positions(0.2 * net["R"], _positions(net))
    v_small = net["site_energies"][N:]
    v_large = np.where(np.arange(P) % 2 == 0, 75.0, -75.0)     # |V~| = 108
    es = lay.energy_sets(np.stack([net["E_constant"]] * 2), np.stack([v_small, v_large]))
    K, hops = 512, 5000
    trajs = np.concatenate([
        engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), hops,
                          net["occupation"], prehops=500),
        engine.make_trajs(K, lay.id, es[1], np.arange(K, 2 * K, dtype=np.uint64), hops,
                          net["occupation"], prehops=500)])
    b = engine.Batch(trajs)
    out = {}
    try:
        for form in (1, 2, 3):
            print(f"[gpu] electrode closed form / fallback, form {form}", flush=True)
            b.run(precision=_lib.FP32_BATCHED, table_form=form)
            b.sync()
            r = b.results()
            assert (r.status == 0).all() and (r.hops_done == hops).all()
            out[form] = r
    finally:
        b.close()
    alone = engine.Batch(trajs[K:])
    try:
        alone.run(precision=_lib.FP32_BATCHED, table_form=1)
        alone.sync()
        a = alone.results()
    finally:
        alone.close()
    assert np.array_equal(a.eo, out[1].eo[K:]) and np.array_equal(a.time, out[1].time[K:])
    assert np.array_equal(a.occ_bits, out[1].occ_bits[K:])
    lay.close()
    for half, sl in (("closed form", slice(0, K)), ("fallback", slice(K, 2 * K))):
        pos = out[3].currents(P)[sl]
        for form in (1, 2):
            _compare(out[form].currents(P)[sl], pos, f"{half}: table {form} vs positions")


@pytest.mark.parametrize("N,P,form", [(30, 8, 1), (30, 8, 3), (50, 8, 2), (90, 10, 2)])
def test_fast_hops_and_leaps_equal_general_steps(gpu_engine, N, P, form):
    """Kernels that record take the general step on every hop; the others take
    fast hops and leaps over runs of hops back.  Both give the same trajectories
    bit for bit: clock, tallies and occupation (S = 38 and 58 one-chunk, S = 100
    two chunks)."""
    net = random_network(40 + N, N, P, I_0=50.0, a=0.1, vmax=1.0)
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    lay.set_positions(0.1 * net["R"], _positions(net))
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    K, hops = 256, 20_000
    b = engine.Batch(engine.make_trajs(K, lay.id, es[0], np.arange(K, dtype=np.uint64), hops,
                                       net["occupation"], prehops=1000))
    print(f"[gpu] fast path vs general steps: S={N + P} form {form}", flush=True)
    out = []
    for rec in (False, True):
        b.run(precision=_lib.FP32_BATCHED, table_form=form, record=rec)
        b.sync()
        r = b.results()
        assert (r.status == 0).all()
        out.append(r)
    b.close()
    lay.close()
    assert np.array_equal(out[0].time, out[1].time)
    assert np.array_equal(out[0].eo, out[1].eo)
    assert np.array_equal(out[0].occ_bits, out[1].occ_bits)
