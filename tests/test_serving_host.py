"""Host logic of the serving loop and the autotuner (no GPU)."""
import types

import numpy as np
import pytest

from kmc_dn_amd import _lib, autotune, serving


def test_popcount_rows():
    bits = np.array([[0, 0, 0, 0], [1, 3, 0, 2**63], [2**64 - 1, 0, 0, 0]], dtype=np.uint64)
    assert list(serving.popcount_rows(bits)) == [0, 4, 64]


def test_bytes_per_trajectory_counts_device_records():
    # descriptor + output record + group entry + E_constant and V doubles
    assert serving.bytes_per_trajectory(30, 8) == 88 + 184 + 16 + 8 * 38


def test_trajectory_seed_is_unique_per_request_and_slot():
    seeds = {serving.trajectory_seed(r, j) for r in range(50) for j in range(300)}
    assert len(seeds) == 50 * 300


def _stub(N, P, occ_counts, chunk=8192, K=1024):
    n = len(occ_counts)
    s = types.SimpleNamespace(chunk=chunk, K=K, REF_COST=serving.Server.REF_COST,
                              round_s=chunk * serving.Server.REF_HOP_S)
    s.lN, s.lP = np.array([N]), np.array([P])
    s.lC = (s.lN + s.lP + 63) // 64
    s.s_lay = np.zeros(n, np.int64)
    occ = np.zeros((n, _lib.OCCW), np.uint64)
    for i, c in enumerate(occ_counts):
        for w in range(_lib.OCCW):
            k = min(64, max(0, c - 64 * w))
            occ[i, w] = np.uint64((1 << k) - 1)
    s.s_occ = occ
    s.s_lat = serving.Server.cost(s, np.arange(n)) * serving.Server.REF_HOP_S / s.REF_COST
    return s


def test_cost_model_grows_with_walked_set_and_chunks():
    s = _stub(100, 8, [0, 10, 50, 90, 100])
    c = serving.Server.cost(s, np.arange(5))
    # the walked set is the smaller of sources and targets: largest at half filling
    assert c[2] >= c[1] and c[2] >= c[3]
    small = serving.Server.cost(_stub(40, 8, [36]), np.arange(1))
    large = serving.Server.cost(_stub(100, 8, [90]), np.arange(1))
    assert large[0] > small[0]


def test_quotas_are_multiples_of_the_recompute_interval():
    s = _stub(100, 8, [0, 10, 50, 90, 100])
    q = serving.Server.quotas(s, np.arange(5))
    assert (q % 1024 == 0).all() and (q >= 1024).all() and (q <= 64 * 8192).all()
    # slower trajectories (measured) get fewer hops for the same round time
    s.s_lat = np.array([1e-6, 2e-6, 4e-6, 1e-5, 1e-3])
    q = serving.Server.quotas(s, np.arange(5))
    assert list(q) == sorted(q, reverse=True) and q[-1] == 1024


def test_autotune_candidates_respect_lds():
    c64 = set(autotune.candidates(64, 56, True))
    assert {(1, 4), (2, 4), (3, 4)} <= c64
    c128 = set(autotune.candidates(128, 120, True))
    assert (1, 1) not in c128 and (2, 1) in c128          # {lt,b}+c table exceeds 160 KiB
    c192 = set(autotune.candidates(192, 180, False))
    assert c192 == set()                                   # no table fits, no positions
    assert autotune.bucket_of(64) == 0 and autotune.bucket_of(65) == 1
    with pytest.raises(ValueError):
        autotune.bucket_of(193)


def test_autotune_cache_roundtrip(tmp_path):
    path = str(tmp_path / "c.json")
    autotune._save_cache(path, {"k": {"form": 3, "waves": 4}})
    assert autotune._load_cache(path) == {"k": {"form": 3, "waves": 4}}
    assert autotune._load_cache(str(tmp_path / "missing.json")) == {}
    assert autotune.cache_key("gfx950", 64, 1.0) != autotune.cache_key("gfx950", 64, 64.0)


def test_merge_plan_fits_the_launch_limit():
    teams = [(64, 1, 1, 4, 900_000), (128, 2, 1, 4, 270_000), (192, 2, 1, 4, 340_000)]
    solos = [(64, 2, 4, 1, 0), (128, 3, 1, 1, 0), (192, 2, 8, 1, 0)]
    sizes = {64: (64, 56), 128: (128, 116), 192: (132, 120)}
    plan = autotune.merge_plan(teams, solos, sizes)
    assert len(plan) <= autotune.MAX_LAUNCHES
    # solo entries merge first into one positions-form entry for every size, then
    # the larger team entries into one table-form kernel whose table still fits
    assert plan == [(64, 1, 1, 4, 900_000), (192, 2, 1, 4, 270_000), (192, 3, 4, 1, 0)]
    assert autotune.table_lds_bytes(2, 132, 120) + 2 * 4 * 3 * 64 * 4 <= autotune.LDS_BYTES
    # first match: every (S, budget) has an entry; long ones land in team entries
    for S in (15, 64, 65, 128, 132):
        for budget in (10_000, 1_000_000):
            hit = [e for e in plan if S <= e[0] and budget >= e[4]][0]
            assert (hit[3] > 1) == (budget >= (900_000 if S <= 64 else 270_000))
    small = autotune.merge_plan(teams[:1], solos[:2], {64: (64, 56), 128: (100, 90)})
    assert small == teams[:1] + solos[:2]
    four = autotune.merge_plan(teams[:2], solos[:2], {64: (64, 56), 128: (108, 100)},
                               max_launches=4)
    assert four == teams[:2] + solos[:2]


def test_split_team_entries():
    """Two table-form team entries are cut where a team workgroup fits twice into
    a CU's LDS; the larger sizes take an entry bounded below by s_min, the
    smaller ones one bounded below by the one-chunk bucket's edge."""
    plan = [(128, 2, 1, 4, 212_420), (192, 2, 1, 4, 256_712), (192, 3, 4, 1, 0)]
    out = autotune.split_team_entries(plan, 132, 120)
    assert out == [(192, 2, 1, 4, 212_420, 0, 93), (92, 2, 1, 4, 212_420, 0, 65),
                   (192, 3, 4, 1, 0)]
    assert autotune.team_lds_bytes(2, 92, 92, 4) <= autotune.LDS_BYTES // 2
    assert autotune.team_lds_bytes(2, 93, 93, 4) > autotune.LDS_BYTES // 2
    assert autotune.team_lds_bytes(2, 132, 120, 4) <= autotune.LDS_BYTES
    # first match with s_min: every (S, budget) has one entry; long ones in teams
    for S in (15, 64, 92, 93, 128, 132):
        for budget in (10_000, 212_420, 1_000_000):
            hit = [e for e in out if (e + (0, 0))[6] <= S <= e[0] and budget >= e[4]][0]
            # one-chunk networks stay in the solo entry (its one-chunk code)
            assert (hit[3] > 1) == (budget >= 212_420 and S > 64)
            assert hit[3] == 1 or (hit[0] == 92) == (S <= 92)
    # unchanged: one team entry, mixed forms, positions teams, a table too large
    assert autotune.split_team_entries(plan[1:], 132, 120) == plan[1:]
    mixed = [(64, 1, 1, 4, 900_000), (192, 2, 1, 4, 270_000), (192, 3, 4, 1, 0)]
    assert autotune.split_team_entries(mixed, 132, 120) == mixed
    pos = [(128, 3, 1, 4, 1000), (192, 3, 1, 4, 2000), (192, 3, 4, 1, 0)]
    assert autotune.split_team_entries(pos, 132, 120) == pos
    assert autotune.split_team_entries(plan, 150, 138) == plan


def test_team_entry_rules():
    """A team entry needs a clear gain against the solo latency inside the
    loaded batch (BATCH_SLOWDOWN times the lone one); its threshold uses that
    in-batch latency and must be reached by some budget of the bucket."""
    lat = {"form": 2, "team": 4, "lat": 2.0e-6, "solo": {"2": 3.1e-6, "3": 6.0e-6}}
    budgets = np.array([1e4, 2e5, 8e5])
    e = autotune.team_entry(192, lat, 3, t_thr=1.5, budgets=budgets)
    thr = int(np.ceil(1.5 / (6.0e-6 * autotune.BATCH_SLOWDOWN)))
    assert e == (192, 2, 1, 4, thr)
    # a team slower alone than a solo wave is inside the batch takes no launch slot
    slow = {"form": 2, "team": 4, "lat": 4.4e-6, "solo": {"1": 1.38e-6, "2": 1.4e-6}}
    assert 4.4e-6 * autotune.TEAM_MIN_GAIN > 1.38e-6 * autotune.BATCH_SLOWDOWN
    assert autotune.team_entry(128, slow, 1, t_thr=1.5, budgets=np.array([1e6])) is None
    # one-chunk buckets take no team entry
    assert autotune.team_entry(64, lat, 3, t_thr=1.5, budgets=budgets) is None
    # a team no faster alone than solo still wins inside the batch: the entry
    # takes teams of TEAM_SIZE waves in the form fastest alone at that size
    even = {"form": 2, "team": 2, "lat": 0.96e-6, "solo": {"2": 0.98e-6, "3": 1.33e-6},
            "teams": {"2/4": 0.99e-6, "3/4": 1.37e-6, "2/2": 0.96e-6}}
    e = autotune.team_entry(192, even, 3, t_thr=1.5, budgets=budgets)
    assert e == (192, 2, 1, 4, int(np.ceil(1.5 / (1.33e-6 * autotune.BATCH_SLOWDOWN))))
    # no budget reaches the threshold
    assert autotune.team_entry(192, lat, 3, t_thr=1.5, budgets=np.array([1e4, 5e4])) is None
    # solo fastest alone
    assert autotune.team_entry(192, {"form": 2, "team": 1, "lat": 3e-6, "solo": {"2": 3e-6}},
                               2, t_thr=1.5, budgets=budgets) is None
    # serving: the threshold is given
    assert autotune.team_entry(192, lat, 3, 0.0, budgets, min_hops=50_000, fixed=True) == \
        (192, 2, 1, 4, 50_000)
