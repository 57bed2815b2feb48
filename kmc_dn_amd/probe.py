"""Measurement probes on the device (python -m kmc_dn_amd.probe {copy,lds}).

copy: a plain device copy of a known byte count -- the reference kernel for
      scaling rocprof's FETCH_SIZE / WRITE_SIZE to bytes;
lds:  LDS read bandwidth of strided ds_read_b32/b64/b128 (--width 4/8/16) and of
      the ds_read_b128 stream with immediate offsets (--width 0, the peak figure);
lone: the hop latency of one trajectory alone (the long tail of an uneven batch
      or of a served request) on a network of N acceptors and P electrodes;
fp64phases: shader cycles per hop and phase of the FP64 dense kernel;
fp32phases: the same for the FP32 batched kernels, on the 4096-trajectory N=30
      ensemble and on lone trajectories (solo and team);
fp32tail: the ensemble's per-trajectory times, and the phases of its slowest,
      median and fastest trajectories rerun alone, with the state store on and off;
chain: latency of dependent FP64 / FP32 adds in one wave.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

from . import _lib, engine


def copy(nbytes: int = 1 << 30, repeats: int = 5) -> dict:
    ms = C.c_double()
    _lib.check(_lib.load().kmc_probe_copy(int(nbytes), int(repeats), C.byref(ms)))
    return {"bytes": nbytes, "ms": ms.value, "GBps": 2 * nbytes / ms.value / 1e6}


def lds(width: int = 16, blocks: int = 0, iters: int = 4096) -> dict:
    if blocks <= 0:
        blocks = 8 * engine.device_info()["cu_count"]
    bps = C.c_double()
    _lib.check(_lib.load().kmc_probe_lds(int(width), int(blocks), int(iters), C.byref(bps)))
    return {"width": width, "blocks": blocks, "GBps": bps.value / 1e9}


def lone(N: int, P: int, form: int, team: int, hops: int = 20_000, seed: int = 1) -> dict:
    """One trajectory of `hops` hops on workloads.network_3d(seed, N, P), run
    twice; the second launch is timed."""
    import numpy as np
    from . import workloads
    net = workloads.network_3d(seed, N, P)
    lay = engine.Layout(N, P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    try:
        lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
        es = lay.energy_sets(net.E_constant, net.site_energies[N:])
        b = engine.Batch(engine.make_trajs(1, lay.id, es[0], np.arange(1, dtype=np.uint64), hops,
                                           net.occupation))
        try:
            for _ in range(2):
                b.run(precision=_lib.FP32_BATCHED, table_form=form, team=team, waves_per_group=1)
                ms = b.sync()
        finally:
            b.close()
    finally:
        lay.close()
    return {"S": N + P, "N": N, "P": P, "form": form, "team": team, "hops": hops,
            "us_per_hop": ms * 1e3 / hops}


FP64_PHASES = ["energies", "pair_lists", "rates", "prefix_chain", "uniforms_clock", "search",
               "hop_records"]


def fp64_phases(hops: int = 20_000, seed: int = 0) -> dict:
    """Shader cycles per hop and phase of the FP64 dense kernel, one trajectory
    alone on the N=30 network of the ensemble workload (the one-trajectory
    launch of the plain path), from its phase-timing instantiation; beside them
    the HIP-event time per hop of the ordinary kernel."""
    import numpy as np
    from . import workloads
    net = workloads.network_2d(seed)
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    try:
        es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])
        b = engine.Batch(engine.make_trajs(1, lay.id, es[0], np.arange(1, dtype=np.uint64), hops,
                                           net.occupation))
        try:
            b.run(precision=_lib.FP64_DENSE)
            b.sync()
            b.run(precision=_lib.FP64_DENSE)
            ms = b.sync()
            cyc = (C.c_uint64 * 8)()
            _lib.check(_lib.load().kmc_probe_fp64_phases(b.id, cyc, 8))
        finally:
            b.close()
    finally:
        lay.close()
    per = [cyc[k] / hops for k in range(len(FP64_PHASES))]
    return {"hops": hops, "us_per_hop_events": ms * 1e3 / hops,
            "cycles_per_hop": dict(zip(FP64_PHASES, [round(x, 1) for x in per])),
            "cycles_per_hop_total": round(sum(per), 1)}


FP32_PHASES = ["uniforms_recompute", "miss_update_row_sums", "miss_scan_insert", "clock_select",
               "partner", "hop", "fast_hops_leaps"]


def _fp32_phase_run(b, run: dict, hops_total: int) -> dict:
    """Phase cycles per hop of batch b (run options `run`) beside the ordinary
    kernel's HIP-event time."""
    for _ in range(2):
        b.run(**run)
        ms = b.sync()
    o = _lib.kmc_run_opts(_lib.FP32_BATCHED, int(run.get("waves_per_group", 0)), 0, 0, 0, 0, 0,
                          int(run.get("table_form", 0)), 0, int(run.get("team", 1)), 0,
                          int(run.get("state_store", 0)))
    cyc = (C.c_uint64 * 11)()
    _lib.check(_lib.load().kmc_probe_fp32_phases(b.id, C.byref(o), cyc, 11))
    per = [cyc[k] / hops_total for k in range(len(FP32_PHASES))]
    return {"kernel_ms": round(ms, 3), "hit_rate": round(cyc[7] / hops_total, 5),
            "partner_hit_rate": round(cyc[8] / hops_total, 5),
            "leap_rate": round(cyc[9] / hops_total, 5),
            "store_restore_rate": round(cyc[10] / hops_total, 6),
            "cycles_per_hop": dict(zip(FP32_PHASES, [round(x, 1) for x in per])),
            "cycles_per_hop_total": round(sum(per), 1)}


def fp32_phases_ensemble(K: int = 4096, hops: int = 100_000) -> dict:
    """Phase cycles per hop and trajectory of the solo table kernel (4 waves
    per workgroup, {lt, lt+c} table) over the K-trajectory N=30 ensemble: every
    wave's cycles, with the waves that share its SIMD interleaved."""
    import numpy as np
    from . import workloads
    ens = workloads.ensemble(0, K, hops)
    net = ens.net
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    try:
        es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])
        b = engine.Batch(engine.make_trajs(K, lay.id, es[0], ens.seeds, hops, net.occupation))
        try:
            r = _fp32_phase_run(b, dict(precision=_lib.FP32_BATCHED, waves_per_group=4,
                                        table_form=1), K * hops)
        finally:
            b.close()
    finally:
        lay.close()
    r.update({"workload": f"{K} x {hops} hops, N=30 P=8", "waves_per_group": 4, "table_form": 1})
    return r


def fp32_phases_lone(N: int, P: int, form: int, team: int, hops: int = 100_000,
                     seed: int = 1) -> dict:
    """Phase cycles per hop of one trajectory alone on workloads.network_3d(seed,
    N, P): solo (4-wave workgroup, one trajectory) or a team of 4."""
    import numpy as np
    from . import workloads
    net = workloads.network_3d(seed, N, P)
    lay = engine.Layout(N, P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    try:
        lay.set_positions(net.ab, np.concatenate([net.acceptors, net.electrodes[:, :3]], 0))
        es = lay.energy_sets(net.E_constant, net.site_energies[N:])
        b = engine.Batch(engine.make_trajs(1, lay.id, es[0], np.arange(1, dtype=np.uint64), hops,
                                           net.occupation))
        try:
            r = _fp32_phase_run(b, dict(precision=_lib.FP32_BATCHED, table_form=form, team=team,
                                        waves_per_group=4 if team <= 1 else 1), hops)
        finally:
            b.close()
    finally:
        lay.close()
    r.update({"S": N + P, "N": N, "P": P, "form": form, "team": team, "hops": hops,
              "us_per_hop": round(r["kernel_ms"] * 1e3 / hops, 3)})
    return r


def fp32_tail(K: int = 4096, hops: int = 100_000, group: int = 32) -> list:
    """The launch's tail: the K-trajectory N=30 ensemble's per-trajectory device
    times (kmc_batch_timing), then the `group` slowest, median and fastest
    trajectories rerun as batches of their own (a SIMD each), with the state
    store on and off: their times and phase cycles per hop."""
    import numpy as np
    from . import workloads
    ens = workloads.ensemble(0, K, hops)
    net = ens.net
    lay = engine.Layout(net.N, net.P, net.nu, net.kT, net.I_0, net.R, net.distances,
                        net.transitions_constant)
    out = []
    try:
        es = lay.energy_sets(net.E_constant, net.site_energies[net.N:])
        b = engine.Batch(engine.make_trajs(K, lay.id, es[0], ens.seeds, hops, net.occupation))
        try:
            b.run(precision=_lib.FP32_BATCHED, waves_per_group=4, table_form=1)
            ms = b.sync()
            sec = b.timing()
        finally:
            b.close()
        order = np.argsort(sec)
        out.append({"workload": f"{K} x {hops} hops, N=30 P=8", "kernel_ms": round(ms, 3),
                    "traj_ms_p0_50_90_99_100": [round(float(x) * 1e3, 3)
                                                for x in np.percentile(sec, [0, 50, 90, 99, 100])]})
        mid = K // 2 - group // 2
        for name, idx in (("slowest", order[-group:]), ("median", order[mid:mid + group]),
                          ("fastest", order[:group])):
            g = engine.Batch(engine.make_trajs(group, lay.id, es[0], ens.seeds[idx], hops,
                                               net.occupation))
            try:
                for store in (0, -1):
                    run = dict(precision=_lib.FP32_BATCHED, waves_per_group=4, table_form=1,
                               state_store=store)
                    g.run(**run)
                    g.sync()
                    t = g.timing()   # the ordinary kernel's, not the phase-timing run's
                    r = _fp32_phase_run(g, run, group * hops)
                    r.update({"group": f"{group} {name}", "state_store": store,
                              "traj_ms_mean_max": [round(float(t.mean()) * 1e3, 3),
                                                   round(float(t.max()) * 1e3, 3)]})
                    out.append(r)
            finally:
                g.close()
    finally:
        lay.close()
    return out


CHAIN_KINDS = ["v_add_f64 dependent (64 lanes)", "v_add_f64 dependent (1 lane)",
               "v_add_f32 dependent", "v_fma_f64 dependent", "v_add_f64 4 interleaved chains"]


def chain_latency(iters: int = 4096) -> dict:
    """Shader cycles per instruction of dependent FP64/FP32 add chains in one
    wave (the FP64 prefix chain's bound) and of four interleaved chains."""
    out = {}
    for k, name in enumerate(CHAIN_KINDS):
        c = C.c_double()
        _lib.check(_lib.load().kmc_probe_chain(k, iters, C.byref(c)))
        out[name] = round(c.value, 2)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "lds", "lone", "fp64phases", "fp32phases", "fp32tail",
                                     "chain"])
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--width", type=int, default=-1, help="4, 8, 16, 0 (stream); -1 all")
    ap.add_argument("--N", type=int, default=30)
    ap.add_argument("--P", type=int, default=8)
    ap.add_argument("--form", type=int, default=1, help="1 {lt, lt+c}, 2 {lt, c}, 3 positions")
    ap.add_argument("--team", type=int, default=1)
    ap.add_argument("--hops", type=int, default=20_000)
    args = ap.parse_args()
    engine.init(0)
    if args.what == "copy":
        print(json.dumps(copy(args.bytes, args.repeats)))
    elif args.what == "lone":
        print(json.dumps(lone(args.N, args.P, args.form, args.team, args.hops)))
    elif args.what == "fp64phases":
        print(json.dumps(fp64_phases(args.hops)))
    elif args.what == "fp32phases":
        print(json.dumps(fp32_phases_ensemble()))
        for N, P, form in [(30, 8, 1), (56, 8, 1), (120, 12, 2), (120, 12, 3)]:
            for team in (1, 4):
                if team == 1 and form == 3 or team == 4 and N + P <= 64:
                    continue
                print(json.dumps(fp32_phases_lone(N, P, form, team)), flush=True)
    elif args.what == "fp32tail":
        for r in fp32_tail():
            print(json.dumps(r), flush=True)
    elif args.what == "chain":
        print(json.dumps(chain_latency()))
    else:
        for w in (4, 8, 16, 0) if args.width < 0 else (args.width,):
            print(json.dumps(lds(w)))


if __name__ == "__main__":
    main()
