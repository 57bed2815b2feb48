"""The oracle's C restatement under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/Makefile's `sanitize` target builds liboracle_san.so from the same
source with -fsanitize=address,undefined -fno-sanitize-recover=all; a child
process preloads libasan, loads that build (KMC_ORACLE_LIB) and drives every
oracle entry point over small networks: the numba and Go loops (record, replayed
uniforms, event logs, the occupation cache, pruning, the OpenMP ensembles), the
mean-field estimators and Philox.  Any out-of-bounds access, use after free or
undefined operation aborts the child with a report.  (The engine's host code
needs a HIP device for all but its argument checks; the GPU tests exercise it.)
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path.insert(0, {root!r})
import numpy as np
import oracle
from tests.netgen import random_network

for seed, (N, P) in enumerate([(10, 2), (30, 8), (70, 11)]):
    net = random_network(seed, N, P, I_0=50.0, a=0.1)
    args = (N, P, 1.0, 1.0, net["I_0"], net["R"], net["distances"], net["E_constant"],
            net["site_energies"], net["transitions_constant"])
    U = np.random.RandomState(seed).rand(2 * 500)
    r = oracle.numba_simulate(*args, net["occupation"].copy(), np.zeros(P, np.int64), 500,
                              record=True, uniforms=U, log=True)
    assert r.status == 0
    r = oracle.numba_simulate(*args, net["occupation"].copy(), np.zeros(P, np.int64), 500,
                              seed=seed, hop0=7)
    for cache in (False, True):
        g = oracle.go_simulate(*args, np.zeros(N, np.uint8), 500, use_cache=cache, record=True,
                               log=True)
        g = oracle.go_simulate(*args, np.zeros(N, np.uint8), 500, use_cache=cache,
                               full_recompute=True, prune_threshold=1e-7, seed=3)
    occ = net["occupation"]
    oracle.numba_ensemble(*args, occ, np.arange(4, dtype=np.uint64), 300, nthreads=2)
    stats = np.zeros((4, 3), np.int64)
    oracle.go_ensemble(*args, np.zeros(N, np.uint8), np.arange(4, dtype=np.uint64), 300,
                       use_cache=True, cache_stats=stats, nthreads=2)
    f = oracle.prob_simulate(*args, occ.astype(np.float64) * 0.5, 50)
    c = oracle.average_current(*args, occ.astype(np.float64) * 0.5)
oracle.philox4x32_10(np.arange(4, dtype=np.uint32), np.arange(2, dtype=np.uint32))
oracle.stream_uniforms(5, 3, 130)
print("sanitized oracle ok")
"""


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True, check=True).stdout.strip()
    if not os.path.isabs(asan):
        pytest.skip("gcc has no libasan runtime")
    env = dict(os.environ, LD_PRELOAD=asan,
               KMC_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_san.so"),
               # the Python interpreter's own allocations are not the oracle's leaks
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               OMP_NUM_THREADS="2")
    p = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "sanitized oracle ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
