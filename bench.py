#!/usr/bin/env python3
"""Benchmark of the KMC hop loop on one MI355X.

Headline: hops/s of the FP32-rate / FP64-accumulate batched kernel on the
4096-trajectory ensemble of the N=30 network (30 acceptors, 3 donors, 8
perimeter electrodes, a = 0.1, I_0 = 50 kT, V ~ U(-1, 1), Laplace potential,
Philox keys 0..4095, 1e5 hops per trajectory).  One step = one launch over the
whole ensemble (4.096e8 hops) with every input already resident in HBM.

Beside it, on the same layout and in the same run:
  * the plain FP64 path (reference arithmetic, one trajectory per launch);
  * the LDS roofline of the batched kernel (algorithmic LDS bytes per hop of
    the reference's dense formulation / kernel time from HIP events);
  * the CPU baseline: the oracle's restatement of the numba loop, ensemble-
    parallel over this host's cores on a bounded sample (plus the Go-loop
    restatement).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "hops/sec, 4096-traj N=30 ensemble vs plain; p50/p99 latency of served IV jobs"
# MI355X: 256 CUs at up to 2.4 GHz; CDNA4's LDS (160 KiB per CU) delivers 256 bytes per
# clock per CU (twice CDNA3's 128) -> 256 * 256 * 2.4e9 B/s = 157.3 TB/s.  The roofline
# divides by the LDS read peak measured in the same run (ds_read_b128 stream probe).
LDS_PEAK_SPEC_GBS = 256 * 256 * 2.4
# HBM bytes per launch of the batched kernel on the default workload (4096 x 1e5 hops,
# N=30/P=8, auto table form and workgroup size): rocprofv3 --pmc FETCH_SIZE 5738.81 KiB
# and WRITE_SIZE 17024 KiB per dispatch, scaled by the copy-kernel calibration factors
# 2.0 (FETCH) and 1.0 (WRITE): (2 * 5738.81 + 17024) * 1024 bytes -- the launch's own
# I/O (1.1 MB) and register-spill scratch (the one-chunk kernels are held to 128 VGPRs).
HBM_TRAFFIC_BYTES = 29_185_659
# VALU and scalar instructions per hop and wave of the batched kernel on the default
# workload (rocprofv3 --pmc SQ_INSTS_VALU 2.561e9, SQ_INSTS_SALU 2.362e9 per dispatch of
# 4.096e8 hops), and the issue peak of each: one wave64 VALU instruction per 4 cycles on
# each of a CU's 4 SIMDs, and one scalar instruction per cycle from the CU's shared
# scalar unit, i.e. one per CU and cycle each, 256 CUs at 2.4 GHz.
VALU_INSTS_PER_HOP = 6.25
SALU_INSTS_PER_HOP = 5.77
ISSUE_PEAK = 256 * 2.4e9


def lds_bytes_per_hop(N: int, P: int) -> int:
    """Algorithmic LDS bytes of one hop in the reference's dense formulation.

    4*S^2 (FP32 log-prefactor per ordered pair) + 4*N^2 (FP32 Coulomb kernel per
    acceptor pair) + 4*S (site energies) + 8*N (two Coulomb columns of the
    rank-1 update); S = N + P.
    """
    S = N + P
    return 4 * S * S + 4 * N * N + 4 * S + 8 * N


def host_cores() -> int:
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baselines(ens, seconds: float) -> tuple[dict, dict, dict]:
    """Oracle restatements timed on this host: the numba loop (FP64), the Go loop
    (FP32) without and with its occupation-state cache (simulation.go:222-296,
    here with a collision-free key), and the cache's hit rate."""
    import oracle
    net = ens.net
    cores = host_cores()
    args = net.oracle_args()
    # calibrate the per-core rate on a short single-thread run
    t0 = time.perf_counter()
    oracle.numba_ensemble(*args, net.occupation, np.arange(1), 1000, nthreads=1)
    per_core = 1000 / (time.perf_counter() - t0)
    ntraj = 4 * cores
    hops = max(1000, int(per_core * seconds / 4))
    seeds = ens.seeds[:ntraj]
    t0 = time.perf_counter()
    _, _, _, st = oracle.numba_ensemble(*args, net.occupation, seeds, hops, nthreads=cores)
    dt = time.perf_counter() - t0
    numba = {"value": ntraj * hops / dt, "unit": "hops/s", "cores": cores, "kind": "port",
             "sample": f"oracle restatement of the numba loop (FP64 dense, "
                       f"kmc_dopant_networks.py:33-135), {ntraj} trajectories x {hops} hops of "
                       f"the same layout, OpenMP one trajectory per thread, {dt:.1f} s",
             "status_ok": bool((st == 0).all())}
    t0 = time.perf_counter()
    occ0 = np.zeros(net.N, np.uint8)
    _, _, _, st = oracle.go_ensemble(*args, occ0, seeds, hops, nthreads=cores)
    dt = time.perf_counter() - t0
    go = {"value": ntraj * hops / dt, "unit": "hops/s", "cores": cores, "kind": "port",
          "sample": f"oracle restatement of the Go loop (FP32, incremental energies, no cache, "
                    f"simulation.go:194-325), {ntraj} x {hops} hops, {dt:.1f} s",
          "status_ok": bool((st == 0).all())}
    # the same loop with the Go library's occupation-state cache (record_problist):
    # a revisited occupation reuses its stored cumulative rate list
    stats = np.zeros((ntraj, 3), np.int64)
    chops = hops * 32
    t0 = time.perf_counter()
    _, _, _, st = oracle.go_ensemble(*args, occ0, seeds, chops, nthreads=cores, use_cache=True,
                                     cache_max_bytes=1 << 26, cache_stats=stats)
    dt = time.perf_counter() - t0
    hit = float(stats[:, 0].sum() / max(1, stats[:, 0].sum() + stats[:, 1].sum()))
    go_cache = {"value": ntraj * chops / dt, "unit": "hops/s", "cores": cores, "kind": "port",
                "sample": f"oracle restatement of the Go loop with the occupation cache "
                          f"(simulation.go:222-296, collision-free key), {ntraj} x {chops} hops "
                          f"from an empty network, {dt:.1f} s",
                "cache_hit_rate": hit,
                "states_per_trajectory": float(stats[:, 2].mean()),
                "status_ok": bool((st == 0).all())}
    return numba, go, go_cache


def fixture_cache_hit_rate(hops: int = 1_000_000, k: int = 4) -> dict:
    """Hit rate of the Go cache on the reference's own +-300 kT fixture networks
    (thesis_indrek/tests, the setting of its published cache timings,
    printTables.py:118-119), go_simulation protocol: empty start."""
    import oracle
    fx = np.load(os.path.join(ROOT, "tests", "golden", "kmc_fixtures.npz"))
    from kmc_dn_amd import tables
    out = {}
    for fam in ("rnd", "xor"):
        stats = np.zeros((k, 3), np.int64)
        N, P = int(fx["N"]), int(fx["P"])
        acc, el = fx[f"{fam}_acceptors"][0], fx[f"{fam}_electrodes"][0]
        d = tables.calc_distances(acc, el)
        tc = tables.calc_transitions_constant(d, float(fx["ab"]), float(fx["nu"]))
        se = np.zeros(N + P)
        se[N:] = el[:, 3]
        t0 = time.perf_counter()
        oracle.go_ensemble(N, P, 1.0, 1.0, float(fx["I_0"]), float(fx["R"]), d,
                           fx[f"{fam}_E_constant"][0], se, tc, np.zeros(N, np.uint8),
                           np.arange(k, dtype=np.uint64), hops, use_cache=True,
                           cache_max_bytes=1 << 26, cache_stats=stats, nthreads=k)
        dt = time.perf_counter() - t0
        out[fam] = {"hit_rate": float(stats[:, 0].sum() / (stats[:, 0].sum() + stats[:, 1].sum())),
                    "states_per_trajectory": float(stats[:, 2].mean()),
                    "hops_per_s_per_core": hops / dt}
    return out


def iv_leg(lay, args) -> dict:
    """IV sweep: electrode 0 over linspace(-2, 2, n), the others held, R replicas each,
    all points x replicas in one launch on the shared layout."""
    from kmc_dn_amd import _lib, iv, workloads
    sw = workloads.iv_sweep(0, args.iv_points, args.iv_replicas, args.iv_hops)
    b, _ = iv.iv_batch(lay, sw.E_sets, sw.V_sets, sw.replicas, sw.hops, sw.net.occupation)
    t0 = time.perf_counter()
    b.run(precision=_lib.FP32_BATCHED)
    ms = b.sync()
    wall = time.perf_counter() - t0
    r = b.results()
    assert (r.status == 0).all()
    mean, se = iv.currents_from(r, args.iv_points, args.iv_replicas, sw.net.P)
    hops = args.iv_points * args.iv_replicas * args.iv_hops
    b.close()
    return {"value": hops / wall, "unit": "hops/s", "kernel_ms": ms,
            "trajectories": args.iv_points * args.iv_replicas, "hops_per_traj": args.iv_hops,
            "workload": f"electrode 0 swept over linspace(-2,2,{args.iv_points}), others held; "
                        f"{args.iv_replicas} replicas per point; energy sets by Laplace-basis "
                        "superposition",
            "current_e0_first_last": [float(mean[0, 0]), float(mean[0, -1])]}


def uneven_leg(args) -> dict:
    """2048 distinct 3D networks (N in [10,120], 4-12 electrodes, log-uniform budgets
    1e4..1e6) in one graph-captured launch under the autotuned bucket plan."""
    from kmc_dn_amd import uneven, workloads
    t0 = time.perf_counter()
    ub = workloads.uneven_batch(seed=0, n=args.uneven_n)
    dev = uneven.upload(ub)
    setup = time.perf_counter() - t0
    try:
        t0 = time.perf_counter()
        from kmc_dn_amd import autotune
        plan = autotune.plan_for(dev.trajs, dev.sizes)
        tune_s = time.perf_counter() - t0
        run = uneven.run(dev, plan=plan, use_graph=True)
        uneven.check_invariants(dev, run.result)
    finally:
        dev.close()
    S = np.array([n.N + n.P for n in ub.nets])
    N = np.array([n.N for n in ub.nets])
    P = np.array([n.P for n in ub.nets])
    bph = np.array([lds_bytes_per_hop(int(a), int(b)) for a, b in zip(N, P)])
    achieved = float(np.sum(bph * ub.hops)) / (run.kernel_ms * 1e-3) / 1e9
    longest = int(np.argmax(ub.hops))
    return {"value": run.hops_per_s, "unit": "hops/s", "kernel_ms": run.kernel_ms,
            "hops": run.hops, "networks": args.uneven_n, "plan": [list(p) for p in run.plan],
            "S_range": [int(S.min()), int(S.max())],
            "longest_budget": int(ub.hops[longest]), "longest_S": int(S[longest]),
            "lds_roofline_frac": achieved / LDS_PEAK_SPEC_GBS,
            "setup_s": setup, "tune_s": tune_s,
            "workload": "2048 seeded 3D networks in 1x1x0.2, N~U{10..120}, M=round(N/10), "
                        "P~U{4..12} side-face electrodes, V~U(-1,1), hops log-uniform [1e4,1e6]; "
                        "one trajectory per network; bucket kernels forked in one hipGraph"}


def iv_job_leg(srv, nets, args) -> dict:
    """Served IV jobs: Poisson arrivals of IV-sweep requests (16 voltages of
    electrode 0 over [-2, 2], the others held at U(-1.5, 1.5), 4 replicas of 2e5
    hops per voltage) over the 64 resident layouts, admitted by device memory and
    run by continuous batching beside one another."""
    from kmc_dn_amd import serving, workloads
    out = {"workload": f"IV jobs on 64 resident 2D layouts N~U{{20..100}}, 8 electrodes: "
                       f"electrode 0 over linspace(-2,2,{args.ivjob_points}), the others "
                       f"U(-1.5,1.5); {args.ivjob_replicas} replicas x {args.ivjob_hops} hops "
                       "per voltage; Poisson arrivals; latency = completion - arrival"}
    for rate in args.ivjob_rates:
        reqs = workloads.iv_request_stream(rate, max(2, int(rate * args.ivjob_seconds)), nets,
                                           seed=100 + int(rate * 10), points=args.ivjob_points,
                                           replicas=args.ivjob_replicas, hops=args.ivjob_hops)
        sreq = [serving.ServeRequest(r.rid, r.t_arrive, r.layout, r.V, r.hops, r.replicas)
                for r in reqs]
        hops = int(sum(r.hops * r.V.shape[0] * r.replicas for r in sreq))
        r0 = srv.rounds
        done, wall = serving.serve_stream(srv, sreq, realtime=True, max_wall=300)
        d = serving.latency_summary(done, wall, hops)
        d.update(offered_req_per_s=rate, requests=len(sreq), rounds=srv.rounds - r0,
                 offered_hops_per_s=hops / max(sreq[-1].t_arrive, 1e-9))
        assert d["failed"] == 0
        out[f"{rate}/s"] = d
    return out


def serving_leg(args) -> dict:
    """Poisson request streams over 64 resident layouts, served by continuous
    batching; beside them the one-launch-per-request baseline, and the served
    IV-job streams on the same server."""
    from kmc_dn_amd import serving, workloads
    nets, bases = workloads.resident_layouts(seed=0, n=64, N_range=(20, 100))
    srv = serving.Server(nets, bases, chunk_hops=args.serve_chunk)
    out = {"plan": [list(p) for p in srv.plan], "chunk_hops": args.serve_chunk,
           "workload": "64 resident 2D layouts N~U{20..100}, 8 electrodes; requests: one "
                       "layout, V~U(-1.5,1.5)^8, hops lognormal(median 5e4, sigma 1), Poisson "
                       "arrivals; latency = completion - arrival"}
    try:
        for rate in args.serve_rates:
            reqs = workloads.request_stream(rate, int(rate * args.serve_seconds), nets, seed=rate)
            sreq = [serving.ServeRequest(r.rid, r.t_arrive, r.layout, r.V, r.hops) for r in reqs]
            hops = int(sum(r.hops for r in sreq))
            r0, h0 = srv.rounds, srv.hops_done
            done, wall = serving.serve_stream(srv, sreq, realtime=True, max_wall=300)
            d = serving.latency_summary(done, wall, hops)
            d.update(offered_req_per_s=rate, requests=len(sreq), rounds=srv.rounds - r0,
                     offered_hops_per_s=hops / sreq[-1].t_arrive,
                     kernel_ms_total=float(np.sum(srv.round_ms[r0:])),
                     round_ms_mean=float(np.mean(srv.round_ms[r0:])),
                     host_s={k: round(v, 3) for k, v in srv.host_s.items()})
            for k in srv.host_s:
                srv.host_s[k] = 0.0
            assert d["failed"] == 0 and srv.hops_done - h0 == hops
            out[f"{rate}/s"] = d
        reqs = workloads.request_stream(args.serve_rates[0], args.serve_baseline_n, nets,
                                        seed=args.serve_rates[0])
        sreq = [serving.ServeRequest(r.rid, r.t_arrive, r.layout, r.V, r.hops) for r in reqs]
        done, wall = serving.serve_one_launch_per_request(srv, sreq, realtime=True)
        d = serving.latency_summary(done, wall, int(sum(r.hops for r in sreq)))
        d.update(offered_req_per_s=args.serve_rates[0], requests=len(sreq),
                 sample=f"first {len(sreq)} requests of the {args.serve_rates[0]}/s stream, "
                        "each launched alone on one stream as it arrives")
        out["one_launch_per_request"] = d
        if not args.no_ivjobs:
            out["iv_jobs"] = iv_job_leg(srv, nets, args)
    finally:
        srv.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ntraj", type=int, default=4096)
    ap.add_argument("--hops", type=int, default=100_000)
    ap.add_argument("--waves-per-group", type=int, default=0)
    ap.add_argument("--recompute-interval", type=int, default=1024)
    ap.add_argument("--plain-launches", type=int, default=4)
    ap.add_argument("--plain-hops", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--iv-points", type=int, default=1024)
    ap.add_argument("--iv-replicas", type=int, default=64)
    ap.add_argument("--iv-hops", type=int, default=200_000)
    ap.add_argument("--no-iv", action="store_true")
    ap.add_argument("--uneven-n", type=int, default=2048)
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--serve-rates", type=int, nargs="+", default=[200, 1000, 5000])
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-chunk", type=int, default=8192)
    ap.add_argument("--serve-baseline-n", type=int, default=48)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--ivjob-rates", type=float, nargs="+", default=[2.0, 5.0, 10.0])
    ap.add_argument("--ivjob-seconds", type=float, default=4.0)
    ap.add_argument("--ivjob-points", type=int, default=16)
    ap.add_argument("--ivjob-replicas", type=int, default=4)
    ap.add_argument("--ivjob-hops", type=int, default=200_000)
    ap.add_argument("--no-ivjobs", action="store_true")
    ap.add_argument("--no-fp64-ensemble", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (WORLD_SIZE={world}) is not part of "
              "this build: trajectories are independent, so N GPUs would each run their own share "
              "of the ensemble with no exchange.", file=sys.stderr)
        return 2

    from kmc_dn_amd import _lib, engine, workloads

    engine.init(0)
    info = engine.device_info()
    ens = workloads.ensemble(seed=0, K=args.ntraj, hops=args.hops)
    net = ens.net
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])
    batch = engine.Batch(engine.make_trajs(args.ntraj, lay.id, es[0], ens.seeds, args.hops,
                                           net.occupation))
    run = dict(precision=_lib.FP32_BATCHED, waves_per_group=args.waves_per_group,
               recompute_interval=args.recompute_interval)
    for _ in range(args.warmup):
        batch.run(**run)
        batch.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run(**run)
        kernel_ms.append(batch.sync())
    wall = time.perf_counter() - t0
    res = batch.results()
    assert (res.status == 0).all() and (res.hops_done == args.hops).all()
    hops_per_step = args.ntraj * args.hops
    value = hops_per_step * args.steps / wall
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    achieved = lds_bytes_per_hop(net.N, net.P) * hops_per_step / avg_kernel_s / 1e9
    from kmc_dn_amd import probe
    lds_peak = probe.lds(0)["GBps"]
    default_workload = (args.ntraj, args.hops, args.waves_per_group) == (4096, 100_000, 0)
    cur = res.currents(net.P)

    # plain FP64 path: one trajectory per launch
    plain = engine.Batch(engine.make_trajs(1, lay.id, es[0], ens.seeds[:1], args.plain_hops,
                                           net.occupation))
    plain.run(precision=_lib.FP64_DENSE)
    plain.sync()
    pms = []
    tp = time.perf_counter()
    for _ in range(args.plain_launches):
        plain.run(precision=_lib.FP64_DENSE)
        pms.append(plain.sync())
    pwall = time.perf_counter() - tp
    plain_value = args.plain_hops * args.plain_launches / pwall

    # the same FP64 kernel over the whole ensemble (one wave per trajectory): the
    # reference arithmetic at full occupancy, and the FP32 currents' error against it
    fp64_ens = None
    if not args.no_fp64_ensemble:
        b64 = engine.Batch(engine.make_trajs(args.ntraj, lay.id, es[0], ens.seeds, args.hops,
                                             net.occupation))
        t64 = time.perf_counter()
        b64.run(precision=_lib.FP64_DENSE)
        ms64 = b64.sync()
        w64 = time.perf_counter() - t64
        r64 = b64.results()
        b64.close()
        assert (r64.status == 0).all()
        c32, c64 = cur, r64.currents(net.P)
        m32, m64 = c32.mean(0), c64.mean(0)
        se = np.sqrt(c32.var(0, ddof=1) / len(c32) + c64.var(0, ddof=1) / len(c64))
        diff = np.abs(m32 - m64)
        z = np.where(se > 0, diff / np.where(se > 0, se, 1), np.where(diff > 0, np.inf, 0))
        resolvable = np.abs(m64) > 10 * se
        rel = diff / np.maximum(np.abs(m64), 1e-300)
        fp64_ens = {"value": hops_per_step / w64, "unit": "hops/s", "kernel_ms": ms64,
                    "sample": f"{args.ntraj} trajectories x {args.hops} hops, FP64 dense kernel, "
                              "one wave per trajectory, one launch"}
        current_error = {
            "vs": "FP64 dense path (the numba loop's arithmetic) on the same Philox streams",
            "max_abs_z": float(z.max()),
            "max_rel_resolvable": float(rel[resolvable].max()) if resolvable.any() else None,
            "resolvable": int(resolvable.sum()), "unresolvable": int((~resolvable).sum()),
            "resolvable_rule": "|I_fp64| > 10 standard errors of the difference",
            "per_electrode_rel": [float(x) for x in rel],
            "per_electrode_z": [float(x) for x in z]}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "hops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded layout (RandomState(0)), Philox trajectory streams",
        "config": {
            "workload": "4096-trajectory ensemble of one 2D network: N=30 acceptors, M=3 donors, "
                        "P=8 perimeter electrodes, a=0.1 (ab=a*R), I_0=50 kT, V~U(-1,1) kT/e, "
                        "P1 Laplace potential, 1e5 hops per trajectory",
            "n_traj": args.ntraj, "hops_per_traj": args.hops, "N": net.N, "P": net.P,
            "kernel": "fp32 rates / fp64 scan, clock and tallies; one wave per trajectory",
            "recompute_interval": args.recompute_interval, "device": info["name"],
        },
        "roofline": {
            "bound": "lds",
            "achieved": achieved,
            "peak": lds_peak,
            "unit": "GB/s",
            "frac": achieved / lds_peak,
            "traffic": HBM_TRAFFIC_BYTES if default_workload else None,
            "peak_spec": LDS_PEAK_SPEC_GBS,
            "frac_spec": achieved / LDS_PEAK_SPEC_GBS,
            "bytes_per_hop": lds_bytes_per_hop(net.N, net.P),
            "hops_per_launch": hops_per_step,
            "kernel_ms": float(np.mean(kernel_ms)),
            "valu_issue_frac": (VALU_INSTS_PER_HOP * hops_per_step / avg_kernel_s / ISSUE_PEAK
                                if default_workload else None),
            "salu_issue_frac": (SALU_INSTS_PER_HOP * hops_per_step / avg_kernel_s / ISSUE_PEAK
                                if default_workload else None),
            "note": "achieved = algorithmic LDS bytes of the reference's dense formulation "
                    "(9,768 B per hop at N=30, P=8) x hops / kernel time; peak = ds_read_b128 "
                    "stream measured in this run.  frac exceeds 1 because 97% of the hops never "
                    "read those tables: a two-entry state cache serves 98.9% of the hops, an "
                    "LDS store of evicted states most of the rest, and leaps take the runs of "
                    "hops between the two cached states a whole uniform block at a time.  The "
                    "launch ends with its slowest trajectories (those that hop among three or "
                    "four states), bound by the latency of their hop chains (LdsUtil 3%, "
                    "VALUBusy 34%, waves waiting on dependencies 41% of their cycles); "
                    "valu/salu_issue_frac = profiled instructions per hop (6.25 VALU, 5.75 "
                    "scalar) x hops / kernel time / (one per CU and cycle)",
        },
        "plain_fp64": {"value": plain_value, "unit": "hops/s",
                       "kernel_ms": float(np.mean(pms)),
                       "sample": f"{args.plain_launches} launches x {args.plain_hops} hops, "
                                 "one trajectory per launch, FP64 dense reference arithmetic"},
        "speedup_vs_plain_fp64": value / plain_value,
        "mean_current": [float(x) for x in cur.mean(0)],
    }
    if fp64_ens is not None:
        out["plain_fp64_ensemble"] = fp64_ens
        out["speedup_vs_fp64_ensemble"] = value / fp64_ens["value"]
        out["current_error"] = current_error
    if not args.no_iv:
        out["iv_sweep"] = iv_leg(lay, args)
    if not args.no_uneven:
        out["uneven_batch"] = uneven_leg(args)
    if not args.no_serve:
        out["serving"] = serving_leg(args)
        out["latency_ms_p50_p99"] = {k: [v.get("p50_ms"), v.get("p99_ms")]
                                     for k, v in out["serving"].items()
                                     if isinstance(v, dict) and "p50_ms" in v}
        for k, v in out["serving"].get("iv_jobs", {}).items():
            if isinstance(v, dict) and "p50_ms" in v:
                out["latency_ms_p50_p99"][f"iv_jobs {k}"] = [v["p50_ms"], v["p99_ms"]]
    if not args.no_cpu_baseline:
        numba, go, go_cache = cpu_baselines(ens, args.cpu_seconds)
        out["cpu_baseline"] = numba
        out["cpu_baseline_go"] = go
        out["cpu_baseline_go_cache"] = go_cache
        out["go_cache_hit_rate_fixtures_300kT"] = fixture_cache_hit_rate()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
