"""ctypes binding of libkmc_dn_amd.so (the C-ABI in include/kmc_dn_amd.h).

The library is loaded from this package directory, as the reference's
goSimulation/pythonBind.py:64 loads libSimulation.so, but once per process and
by absolute path.  There is no fallback: if the library is missing or fails to
load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmc_dn_amd.so")

NMAX = 256
PMAX = 16
OCCW = 4

(OK, ERR_ARGS, ERR_ZERO_RATE, ERR_NOMEM, ERR_DEVICE, ERR_UNSUPPORTED, ERR_LIMIT, ERR_TIMEOUT,
 ERR_CANCELLED) = range(9)
FP64_DENSE, FP32_BATCHED = 0, 1


class KmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kmc_dn_amd error {code}: {msg}")
        self.code = code


class kmc_traj(C.Structure):
    _fields_ = [("layout", C.c_int32), ("eset", C.c_int32), ("seed", C.c_uint64),
                ("hop0", C.c_uint64), ("prehops", C.c_int64), ("hops", C.c_int64),
                ("occ", C.c_uint64 * OCCW)]


class kmc_run_opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("waves_per_group", C.c_int32),
                ("recompute_interval", C.c_int32), ("record", C.c_int32),
                ("post_hop_record", C.c_int32), ("grid_limit", C.c_int32),
                ("use_graph", C.c_int32), ("table_form", C.c_int32),
                ("resume", C.c_int32), ("team", C.c_int32), ("time_limit_us", C.c_int32),
                ("state_store", C.c_int32)]


class kmc_bucket(C.Structure):
    _fields_ = [("s_max", C.c_int32), ("table_form", C.c_int32),
                ("waves_per_group", C.c_int32), ("team", C.c_int32),
                ("min_hops", C.c_int64), ("grid", C.c_int32), ("s_min", C.c_int32)]


class GoSlice(C.Structure):
    """cgo slice header passed by value (goSimulation/pythonBind.py:28-30)."""
    _fields_ = [("data", C.POINTER(C.c_double)), ("len", C.c_longlong), ("cap", C.c_longlong)]


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)
_i32 = C.c_int32
_i64 = C.c_int64

# name -> (restype, argtypes)
SIGNATURES = {
    "kmc_last_error": (C.c_char_p, []),
    "kmc_version": (C.c_int, []),
    "kmc_init": (C.c_int, [C.c_int]),
    "kmc_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "kmc_reset": (C.c_int, []),
    "kmc_simulate_discrete": (C.c_int, [
        _i64, _i64, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _u8p,
        _i64p, _i64, C.c_int, _dp, _dp, _dp, C.c_uint64, C.c_uint64, _i32p, _dp, _dp]),
    "kmc_simulate_discrete_noalloc": (C.c_int, [
        _i64, _i64, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _u8p,
        _i64p, _i64, C.c_int, _dp, _dp, _dp, C.c_uint64, C.c_uint64, _i32p, _dp, _dp]),
    "kmc_meanfield_simulate": (C.c_int, [_i32, _i32p, _i32p, _i64, _dp, _dp, _dp, _dp, _dp, _dp]),
    "kmc_average_current": (C.c_int, [_i32, _i32p, _i32p, _dp, _dp]),
    "kmc_layout_create": (C.c_int, [_i64, _i64, C.c_double, C.c_double, C.c_double, C.c_double,
                                    _dp, _dp, C.c_double, _i32p]),
    "kmc_layout_destroy": (C.c_int, [_i32]),
    "kmc_layout_set_positions": (C.c_int, [_i32, C.c_double, _dp]),
    "kmc_batch_update": (C.c_int, [_i32, C.POINTER(kmc_traj), _i64]),
    "kmc_batch_set_state": (C.c_int, [_i32, _i64p, _i64p, _dp, _i64p, _u64p]),
    "kmc_eset_create": (C.c_int, [_i32, _i64, _dp, _dp, _i32p]),
    "kmc_eset_destroy": (C.c_int, [_i32]),
    "kmc_eset_write": (C.c_int, [_i32, _dp, _dp]),
    "kmc_eset_write_many": (C.c_int, [_i64, _i32p, _dp, _dp]),
    "kmc_batch_set_buckets": (C.c_int, [_i32, C.POINTER(kmc_bucket), _i32]),
    "kmc_batch_create": (C.c_int, [C.POINTER(kmc_traj), _i64, _i32p]),
    "kmc_batch_run": (C.c_int, [_i32, C.POINTER(kmc_run_opts)]),
    "kmc_batch_sync": (C.c_int, [_i32, _dp]),
    "kmc_batch_wait": (C.c_int, [_i32, C.c_double, _dp]),
    "kmc_batch_results": (C.c_int, [_i32, _dp, _i64p, _u64p, _i32p, _i64p, _dp]),
    "kmc_batch_traffic": (C.c_int, [_i32, _dp, _i64p]),
    "kmc_batch_timing": (C.c_int, [_i32, _dp]),
    "kmc_batch_destroy": (C.c_int, [_i32]),
    "kmc_probe_copy": (C.c_int, [_i64, C.c_int, _dp]),
    "kmc_probe_lds": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp]),
    "kmc_probe_fp64_phases": (C.c_int, [_i32, _u64p, _i32]),
    "kmc_probe_fp32_phases": (C.c_int, [_i32, C.POINTER(kmc_run_opts), _u64p, _i32]),
    "kmc_probe_chain": (C.c_int, [C.c_int, C.c_int, _dp]),
}

_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: build it with "
                                  "`python -m kmc_dn_amd.build` (hipcc, gfx950)")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = load().kmc_last_error()
        raise KmcError(rc, msg.decode() if msg else "")


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)
