"""Serving loop: continuous batching gives each request the result it gets alone."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, serving, workloads

pytestmark = pytest.mark.gpu

# fixed so the test does not time candidates; budgets >= 8000 run in team buckets
PLAN = [(64, 1, 1, 4, 8000), (64, 1, 4, 1, 0), (128, 3, 1, 2, 8000), (128, 3, 4, 1, 0)]


@pytest.fixture(scope="module")
def resident(gpu_engine):
    nets, bases = workloads.resident_layouts(seed=2, n=6, N_range=(20, 100))
    return nets, bases


def _stream(nets, n, rate, seed, median_hops=6000.0):
    reqs = workloads.request_stream(rate, n, nets, seed=seed, median_hops=median_hops,
                                    max_hops=60_000)
    return [serving.ServeRequest(r.rid, r.t_arrive, r.layout, r.V, r.hops) for r in reqs]


def _by_rid(done):
    return {c.rid: c for c in done}


def _assert_same(a, b):
    assert a.status == b.status == 0
    assert np.array_equal(a.eo, b.eo)
    assert np.array_equal(a.time, b.time)
    assert np.array_equal(a.occ_bits, b.occ_bits)


def test_alone_equals_mix(resident):
    nets, bases = resident
    srv = serving.Server(nets, bases, chunk_hops=2048, plan=PLAN)
    try:
        reqs = _stream(nets, 40, 2000.0, seed=11)
        print("[gpu] serving: continuous batching, 40 requests in real time", flush=True)
        mix, wall = serving.serve_stream(srv, reqs, realtime=True, max_wall=60)
        assert len(mix) == 40 and srv.rounds > 3
        print(f"[gpu] serving: {srv.rounds} rounds, {wall:.2f} s", flush=True)
        print("[gpu] serving: the same requests one launch each", flush=True)
        alone, _ = serving.serve_one_launch_per_request(srv, reqs, realtime=False)
        m, a = _by_rid(mix), _by_rid(alone)
        for rid in a:
            _assert_same(m[rid], a[rid])
        # and the chunked path equals one unchunked batch of everything
        print("[gpu] serving: all requests submitted at once", flush=True)
        srv2 = serving.Server(nets, bases, chunk_hops=8192, plan=PLAN)
        try:
            burst, _ = serving.serve_stream(srv2, reqs, realtime=False, max_wall=60)
        finally:
            srv2.close()
        b = _by_rid(burst)
        for rid in a:
            _assert_same(b[rid], a[rid])
    finally:
        srv.close()


def test_iv_requests_and_memory_admission(resident):
    nets, bases = resident
    per = serving.bytes_per_trajectory(nets[0].N, nets[0].P)
    budget = 10 * per          # two IV requests of 3 points x 2 replicas never fit together
    srv = serving.Server(nets, bases, chunk_hops=1024, mem_budget=budget, plan=PLAN)
    peak = []
    orig = srv._admit

    def admit(now):
        orig(now)
        peak.append(srv.mem_used)
    srv._admit = admit
    try:
        rng = np.random.RandomState(4)
        reqs = [serving.ServeRequest(i, 0.0, 0, rng.uniform(-1.5, 1.5, (3, nets[0].P)), 4096,
                                     replicas=2) for i in range(4)]
        print("[gpu] serving: IV requests under a memory budget", flush=True)
        done, _ = serving.serve_stream(srv, reqs, realtime=False, max_wall=60)
        assert max(peak) <= budget and max(peak) == 6 * per
        assert sorted(c.rid for c in done) == [0, 1, 2, 3]
        alone, _ = serving.serve_one_launch_per_request(srv, reqs, realtime=False)
        m, a = _by_rid(done), _by_rid(alone)
        for rid in a:
            assert m[rid].currents.shape == (3, 2, nets[0].P)
            _assert_same(m[rid], a[rid])
        # the two replicas of a point are different streams
        assert not np.array_equal(m[0].eo[0, 0], m[0].eo[0, 1]) or \
            m[0].time[0, 0] != m[0].time[0, 1]
    finally:
        srv.close()


def test_poison_requests_fail_alone(resident):
    nets, bases = resident
    dead = workloads.network_2d(9, N=20, M=2, a=1e-5)     # every rate underflows to zero
    srv = serving.Server(list(nets) + [dead],
                         list(bases) + [workloads.laplace.LaplaceBasis(dead.acceptors,
                                                                       dead.electrodes)],
                         chunk_hops=2048, plan=PLAN)
    try:
        good = _stream(nets, 12, 1e4, seed=5)
        bad_v = np.full(8, np.nan)
        reqs = good[:6] + [serving.ServeRequest(100, 0.0, 1, bad_v, 5000),
                           serving.ServeRequest(101, 0.0, len(nets), np.zeros(8), 5000),
                           serving.ServeRequest(102, 0.0, 99, np.zeros(8), 5000)] + good[6:]
        for r in reqs:
            r.t_arrive = 0.0
        print("[gpu] serving: poison requests in the mix", flush=True)
        done, _ = serving.serve_stream(srv, reqs, realtime=False, max_wall=60)
        d = _by_rid(done)
        assert d[100].status == _lib.ERR_ARGS
        assert d[101].status == _lib.ERR_ZERO_RATE
        assert d[102].status == _lib.ERR_ARGS
        alone, _ = serving.serve_one_launch_per_request(srv, good, realtime=False)
        for c in alone:
            _assert_same(d[c.rid], c)
    finally:
        srv.close()


def test_deadlines_and_cancellation_fail_alone(resident):
    """A request past its deadline, one cancelled while it runs and one cancelled
    before it runs complete with ERR_CANCELLED (the running ones with their
    tallies so far); the others get the results they get alone.  A solo plan, so
    the rounds end at their device time limit."""
    nets, bases = resident
    solo = [(64, 1, 4, 1, 0), (128, 3, 4, 1, 0)]
    srv = serving.Server(nets, bases, chunk_hops=2048, plan=solo)
    try:
        assert srv.time_limit_us > 0
        good = _stream(nets, 10, 1e4, seed=7)
        V = np.zeros(8)
        late = serving.ServeRequest(200, 0.0, 0, V, 3_000_000, deadline_s=0.05)
        victim = serving.ServeRequest(201, 0.0, 1, V, 3_000_000)
        never = serving.ServeRequest(202, 0.0, 2, V, 5000)
        for r in good:
            r.t_arrive = 0.0
        t0 = __import__("time").perf_counter()

        def clock():
            return __import__("time").perf_counter() - t0

        for r in good + [late, victim]:
            srv.submit(r)
        print("[gpu] serving: deadline and cancellation", flush=True)
        srv.round(clock)
        srv.cancel(201)
        srv.cancel(202)
        srv.submit(never)
        while srv.busy():
            srv.round(clock)
            assert clock() < 60
        d = _by_rid(srv.completed)
        for rid in (200, 201, 202):
            assert d[rid].status == _lib.ERR_CANCELLED, (rid, d[rid].message)
        assert (d[200].time > 0).all() and (d[201].time > 0).all()
        assert (d[202].time == 0).all()
        assert srv.mem_used == 0 and not srv.active
        alone, _ = serving.serve_one_launch_per_request(srv, good, realtime=False)
        for c in alone:
            _assert_same(d[c.rid], c)
    finally:
        srv.close()
