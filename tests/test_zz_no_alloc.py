"""kmc_simulate_discrete_noalloc against kmc_simulate_discrete, through the C-ABI.

The allocation-free variant runs the same FP64 kernel with every device buffer
in its static scratch array; on the same inputs every output must be the same
bits: two calls in a row, the second on a smaller network with fewer hops
(the scratch keeps the first call's contents), each with replayed uniforms and
event logs, then on the Philox stream.  Sizes beyond the scratch are refused.
"""
import ctypes as C

import numpy as np
import pytest

from kmc_dn_amd import _lib
from tests.netgen import random_network

pytestmark = pytest.mark.gpu


def _call(name, net, hops, uniforms, seed):
    lib = _lib.load()
    N, P = net["N"], net["P"]
    S = N + P
    se = np.ascontiguousarray(net["site_energies"], dtype=np.float64).copy()
    occ = np.ascontiguousarray(net["occupation"], dtype=np.uint8).copy()
    eo = np.arange(P, dtype=np.int64)   # accumulated into, as the reference does
    traffic = np.zeros((S, S))
    oit = np.zeros(N)
    ev = np.zeros(hops, np.int32)
    tot = np.zeros(hops)
    t = C.c_double()
    u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64)
    p = _lib.ptr
    rc = getattr(lib, name)(
        N, P, 1.0, 1.0, net["I_0"], net["R"], p(np.ascontiguousarray(net["distances"]), _lib._dp),
        p(np.ascontiguousarray(net["E_constant"]), _lib._dp), p(se, _lib._dp),
        p(np.ascontiguousarray(net["transitions_constant"]), _lib._dp), p(occ, _lib._u8p),
        p(eo, _lib._i64p), hops, 1, p(traffic, _lib._dp), p(oit, _lib._dp), p(u, _lib._dp),
        seed, 7, p(ev, _lib._i32p), p(tot, _lib._dp), C.byref(t))
    _lib.check(rc)
    return dict(se=se, occ=occ, eo=eo, traffic=traffic, oit=oit, ev=ev, tot=tot, t=t.value)


def test_noalloc_equals_simulate_discrete(gpu_engine):
    big = random_network(71, 60, 12, I_0=50.0, a=0.1, vmax=1.5)    # S = 72
    small = random_network(72, 20, 6, I_0=30.0, a=0.3, vmax=2.0)   # S = 26
    for label, net, hops in (("S=72", big, 6000), ("S=26", small, 2500)):
        for replay in (True, False):
            u = np.random.RandomState(hops).rand(2 * hops) if replay else None
            print(f"[gpu] noalloc vs simulate_discrete {label} {hops} hops replay={replay}",
                  flush=True)
            a = _call("kmc_simulate_discrete", net, hops, u, 11)
            b = _call("kmc_simulate_discrete_noalloc", net, hops, u, 11)
            for k in a:
                assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (label, replay, k)
            assert a["t"] > 0 and np.abs(a["traffic"]).sum() == hops


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu_engine):
    lib = _lib.load()
    net = random_network(73, 90, 4, I_0=30.0, a=0.3, vmax=1.0)     # S = 94 > 80
    with pytest.raises(_lib.KmcError) as e:
        _call("kmc_simulate_discrete_noalloc", net, 10, None, 1)
    assert e.value.code == _lib.ERR_LIMIT
    small = random_network(74, 10, 4, I_0=30.0, a=0.3, vmax=1.0)
    hops = 65537
    with pytest.raises(_lib.KmcError) as e:
        _call("kmc_simulate_discrete_noalloc", small, hops, np.full(2 * hops, 0.5), 1)
    assert e.value.code == _lib.ERR_LIMIT
    assert b"65536" in lib.kmc_last_error()
