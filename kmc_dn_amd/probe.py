"""Measurement probes on the device (python -m kmc_dn_amd.probe {copy,lds}).

copy: a plain device copy of a known byte count -- the reference kernel for
      scaling rocprof's FETCH_SIZE / WRITE_SIZE to bytes;
lds:  LDS read bandwidth of strided ds_read_b32/b64/b128 (--width 4/8/16) and of
      the ds_read_b128 stream with immediate offsets (--width 0, the peak figure).
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


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "lds"])
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--width", type=int, default=-1, help="4, 8, 16, 0 (stream); -1 all")
    args = ap.parse_args()
    engine.init(0)
    if args.what == "copy":
        print(json.dumps(copy(args.bytes, args.repeats)))
    else:
        for w in (4, 8, 16, 0) if args.width < 0 else (args.width,):
            print(json.dumps(lds(w)))


if __name__ == "__main__":
    main()
