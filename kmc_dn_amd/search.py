"""Search evaluation on the serving loop.

The reference's searches score a candidate by setting electrode voltages per
test, re-solving the potential (update_V) and running go_simulation once per
test:

* dn_search.evaluate_error (dn_search.py:98-119) with average_cumulative_error
  (:136-148): mean excess of |current[index] - expected| over expected_error;
* voltage_search.evaluate_error_corr (voltage_search.py:107-131): separation of
  the output current between true and false tests, weighted by its correlation
  with the ideal output when the classes overlap;
* voltage_search.evaluate_error_diff (:88-105): the separation alone, with the
  input electrodes at the candidate's true_voltage or 0;
* voltage_search.parallel_simulation / evaluate_error_corr_parallel
  (:133-185): the same error over a population simulated together.

Here a candidate is one request of k = len(tests) voltage vectors on a resident
layout (E_constant from the layout's unit-voltage Laplace basis instead of a
solve per test), a population is submitted at once and served by continuous
batching, and the error functions are evaluated on the returned currents.  A
candidate's currents do not depend on the rest of the population: its
trajectories' Philox keys come from its request id.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import serving


def average_cumulative_error(diffs, expected_error: float) -> float:
    """dn_search.average_cumulative_error (dn_search.py:136-148)."""
    diffs = np.asarray(diffs, dtype=np.float64)
    if diffs.size == 0:
        raise ValueError("no expected currents to compare")
    excess = diffs[diffs > expected_error] - expected_error
    return float(excess.sum() / diffs.size)


def dn_search_error(currents, tests, expected_error: float, use_tests: int | None = None) -> float:
    """dn_search.evaluate_error (dn_search.py:98-119) on precomputed currents:
    currents[j] is the (P,) current vector of test j; tests[j] = (electrode
    voltages, [(electrode index, expected current), ...])."""
    n = len(tests) if use_tests is None else use_tests
    diffs = [abs(float(currents[j][index]) - current)
             for j in range(n) for index, current in tests[j][1]]
    return average_cumulative_error(diffs, expected_error)


def perfect_correlation(tests) -> np.ndarray:
    """The ideal output pattern: 10 for tests expecting True, 0 otherwise
    (voltage_search.py:43-48)."""
    return np.array([10 if t[1] else 0 for t in tests])


def _separation(values, tests) -> float:
    lowest_true, highest_false = 1.0, -1.0
    for v, t in zip(values, tests):
        if t[1]:
            lowest_true = min(lowest_true, float(v))
        else:
            highest_false = max(highest_false, float(v))
    return highest_false - lowest_true


def corr_error(values, tests, corr_pow: float = 1.0) -> float:
    """voltage_search.evaluate_error_corr (voltage_search.py:107-131) on the
    output-electrode currents `values` of the tests: the separation when the
    classes overlap (> 0), else corr^corr_pow times it, with corr clipped at 0."""
    values = np.asarray(values, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        corr = np.corrcoef(perfect_correlation(tests), values)[0][1]
    sep = _separation(values, tests)
    if corr < 0:
        corr = 0.0
    if sep > 0:
        return sep
    return float(corr ** corr_pow * sep)


def diff_error(values, tests) -> float:
    """voltage_search.evaluate_error_diff (voltage_search.py:88-105)."""
    return _separation(values, tests)


def voltages_per_test(V, tests, mode: str = "corr", true_voltage: float = 0.0) -> np.ndarray:
    """(len(tests), P) electrode voltages of a candidate with voltages V (P,):
    the input electrodes i < len(test[0]) take test[0][i] ("corr", "dn") or
    true_voltage / 0 by the truth of test[0][i] ("diff"); the others keep V."""
    V = np.asarray(V, dtype=np.float64)
    out = np.repeat(V[None], len(tests), 0)
    for j, t in enumerate(tests):
        for i, x in enumerate(t[0]):
            if mode == "diff":
                out[j, i] = true_voltage if x else 0.0
            else:
                out[j, i] = x
    return out


@dataclass
class Candidate:
    layout: int                  # index of a resident layout of the server
    V: np.ndarray                # (P,) electrode voltages (the inputs are overwritten per test)
    true_voltage: float = 0.0    # input level of "diff" tests


@dataclass
class Evaluation:
    errors: np.ndarray           # (n,) error per candidate
    currents: np.ndarray         # (n, k, P) mean current over replicas per test
    status: np.ndarray           # (n,) 0 or the first failing trajectory's code


def evaluate_population(server: serving.Server, candidates, tests, hops: int,
                        mode: str = "corr", output_electrode: int = -1, corr_pow: float = 1.0,
                        expected_error: float = 0.0, replicas: int = 1, rid0: int = 0,
                        alone: bool = False) -> Evaluation:
    """Errors of a population, all candidates served together (alone=True: one
    launch per candidate, the single-request path).  mode: "corr" / "diff"
    (voltage_search, output current of `output_electrode`) or "dn" (dn_search,
    expected currents per test)."""
    if mode not in ("corr", "diff", "dn"):
        raise ValueError("mode must be 'corr', 'diff' or 'dn'")
    reqs = []
    for n, c in enumerate(candidates):
        Vt = voltages_per_test(c.V, tests, mode, c.true_voltage)
        reqs.append(serving.ServeRequest(rid0 + n, 0.0, c.layout, Vt, int(hops), replicas=replicas))
    if alone:
        done, _ = serving.serve_one_launch_per_request(server, reqs, realtime=False)
    else:
        done, _ = serving.serve_stream(server, reqs, realtime=False)
    by = {d.rid: d for d in done}
    P = np.array([server.nets[c.layout].P for c in candidates])
    cur = np.full((len(candidates), len(tests), P.max(initial=0)), np.nan)
    err = np.full(len(candidates), np.nan)
    st = np.zeros(len(candidates), np.int32)
    for n, c in enumerate(candidates):
        d = by[rid0 + n]
        st[n] = d.status
        if d.status != 0:
            continue
        mean = d.currents.mean(axis=1)                     # (k, P) over replicas
        cur[n, :, :mean.shape[1]] = mean
        if mode == "dn":
            err[n] = dn_search_error(mean, tests, expected_error)
        else:
            out = mean[:, output_electrode]
            err[n] = corr_error(out, tests, corr_pow) if mode == "corr" else diff_error(out, tests)
    return Evaluation(err, cur, st)
