"""Measurement probes on the device (python -m kmc_dn_amd.probe {copy,lds}).

copy: a plain device copy of a known byte count -- the reference kernel for
      scaling rocprof's FETCH_SIZE / WRITE_SIZE to bytes;
lds:  LDS read bandwidth of strided ds_read_b32/b64/b128 (--width 4/8/16) and of
      the ds_read_b128 stream with immediate offsets (--width 0, the peak figure);
lone: the hop latency of one trajectory alone (the long tail of an uneven batch
      or of a served request) on a network of N acceptors and P electrodes.
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


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "lds", "lone"])
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
    else:
        for w in (4, 8, 16, 0) if args.width < 0 else (args.width,):
            print(json.dumps(lds(w)))


if __name__ == "__main__":
    main()
