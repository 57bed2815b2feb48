"""Search evaluation (voltage_search / dn_search error functions) over a served
population equals the single-request path, candidate by candidate."""
import numpy as np
import pytest

from kmc_dn_amd import search, serving, workloads

pytestmark = pytest.mark.gpu

XOR = [((0.0, 0.0), False), ((1.5, 0.0), True), ((0.0, 1.5), True), ((1.5, 1.5), False)]
OUT = 2    # an electrode beside both inputs, which carries current in every test


@pytest.fixture(scope="module")
def server(gpu_engine):
    nets, bases = workloads.resident_layouts(seed=3, n=4, N_range=(20, 50))
    srv = serving.Server(nets, bases, chunk_hops=2048, plan=[(64, 1, 4, 1, 0), (128, 3, 4, 1, 0)])
    yield srv
    srv.close()


def _population(n, seed):
    rng = np.random.RandomState(seed)
    return [search.Candidate(int(rng.randint(4)), rng.uniform(-1.0, 1.0, 8),
                             float(rng.uniform(0.25, 1.0))) for _ in range(n)]


@pytest.mark.parametrize("mode", ["corr", "diff"])
def test_population_equals_single_requests(server, mode):
    cands = _population(10, seed=7)
    print(f"[gpu] search: population of 10 candidates, mode {mode}", flush=True)
    pop = search.evaluate_population(server, cands, XOR, hops=8000, mode=mode, replicas=2,
                                     rid0=1000, output_electrode=OUT)
    print("[gpu] search: the same candidates one launch each", flush=True)
    one = search.evaluate_population(server, cands, XOR, hops=8000, mode=mode, replicas=2,
                                     rid0=1000, output_electrode=OUT, alone=True)
    assert (pop.status == 0).all() and (one.status == 0).all()
    assert np.array_equal(pop.currents, one.currents)
    # NaN where an output current is the same in every test (np.corrcoef's
    # 0/0, as in the reference)
    assert np.array_equal(pop.errors, one.errors, equal_nan=True)
    assert np.isfinite(pop.errors).sum() >= len(cands) // 2
    # the error is the reference's function of the output currents
    for n in range(len(cands)):
        out = pop.currents[n, :, OUT]
        ref = search.corr_error(out, XOR) if mode == "corr" else search.diff_error(out, XOR)
        assert np.array_equal(pop.errors[n], ref, equal_nan=True)


def test_dn_search_error_of_its_own_currents_is_zero(server):
    """dn_search tests whose expected currents are a candidate's served currents
    give that candidate error 0 on a re-evaluation (same request ids, same bits)."""
    cands = _population(4, seed=9)
    tests0 = [((0.3, -0.2), [(7, 0.0)]), ((-0.5, 0.4), [(7, 0.0)])]
    print("[gpu] search: dn_search evaluation", flush=True)
    first = search.evaluate_population(server, cands[:1], tests0, hops=4000, mode="dn", rid0=50)
    cur = first.currents[0]
    tests = [(t[0], [(e, float(cur[j, e])) for e in range(8)]) for j, t in enumerate(tests0)]
    again = search.evaluate_population(server, cands[:1], tests, hops=4000, mode="dn", rid0=50)
    assert again.errors[0] == 0.0
    other = search.evaluate_population(server, cands[:1], tests, hops=4000, mode="dn", rid0=51)
    assert other.errors[0] >= 0.0 and not np.array_equal(other.currents, again.currents)
