"""Pure-Python restatement of the numba hop loop for tiny cases (oracle; test-only).

Follows kmc_dopant_networks.py:33-163 statement by statement with Python
floats (IEEE binary64, no contraction), so it is an independent second reading
of the reference against which the C oracle is checked bit-for-bit on small
networks.  ``math.exp``/``math.log`` are the platform libm, as numba's are.
"""
from __future__ import annotations

import math


def transition_possible(i, j, N, occupation):
    """kmc_dopant_networks.py:137-163."""
    if i >= N and j >= N:
        return False
    elif i >= N:
        return not occupation[j]
    elif j >= N:
        return bool(occupation[i])
    elif i == j:
        return False
    else:
        return bool(occupation[i]) and not occupation[j]


def simulate(N, P, nu, kT, I_0, R, distances, E_constant, site_energies, transitions_constant,
             occupation, electrode_occupation, hops, uniforms, record=False):
    """Returns (time, occupation, electrode_occupation, traffic, occupations_in_time, events).

    ``distances``/``transitions_constant`` are nested lists (S x S);
    ``uniforms`` is a flat list, 2 per hop (clock, event).
    """
    S = N + P
    occupation = [bool(x) for x in occupation]
    eo = [int(x) for x in electrode_occupation]
    se = [float(x) for x in site_energies]
    traffic = [[0.0] * S for _ in range(S)]
    oit = [0.0] * N
    time = 0.0
    events = []
    for hop in range(hops):
        for i in range(N):
            acc = 0.0
            se[i] = E_constant[i]
            for j in range(N):
                if j != i:
                    acc += (1 - int(occupation[j])) / distances[i][j]
            se[i] += -I_0 * R * acc
        trans = [[0.0] * S for _ in range(S)]
        for i in range(S):
            for j in range(S):
                if not transition_possible(i, j, N, occupation):
                    t = 0.0
                else:
                    if i < N and j < N:
                        dE = se[j] - se[i] - I_0 * R / distances[i][j]
                    else:
                        dE = se[j] - se[i]
                    t = nu * math.exp(-dE / kT) if dE > 0 else nu
                trans[i][j] = transitions_constant[i][j] * t
        prob = [0.0] * (S * S)
        for i in range(S):
            for j in range(S):
                k = S * i + j
                prob[k] = trans[i][j] if k == 0 else trans[i][j] + prob[k - 1]
        total = prob[-1]
        scale = 1 / total
        hop_time = -math.log(1.0 - uniforms[2 * hop]) * scale
        u = uniforms[2 * hop + 1]
        event = None
        for k in range(S * S):
            if prob[k] / total >= u:
                event = k
                break
        if event is None:
            event = 0
        frm, to = int(event / S), int(event % S)
        events.append(event)
        if frm < N:
            occupation[frm] = False
        else:
            eo[frm - N] -= 1
        if to < N:
            occupation[to] = True
        else:
            eo[to - N] += 1
        if record:
            traffic[frm][to] += 1
            for i in range(N):
                if occupation[i]:
                    oit[i] += hop_time
        time += hop_time
    return time, occupation, eo, traffic, oit, events


def _fdiv(a, b):
    """IEEE binary64 a / b, including b == 0 (Go and C give +-inf or NaN; Python raises)."""
    if b != 0:
        return a / b
    if a == 0 or math.isnan(a):
        return math.nan
    return math.copysign(math.inf, a) * math.copysign(1.0, b)


def prob_possible(i, j, n_sites, occupation):
    """probTransitionPossible, goSimulation/probabilitySimulation.go:40-50."""
    if i >= n_sites and j >= n_sites:
        return 0.0
    if i >= n_sites:
        return 1 - occupation[j]
    if j >= n_sites:
        return occupation[i]
    return (1 - occupation[j]) * occupation[i]


def prob_simulate(n_sites, n_el, nu, kT, I_0, R, occupation, distances, E_constant,
                  transitions_constant, site_energies, hops, record=False):
    """probSimulate (probabilitySimulation.go:53-157) with calcProbTransitions
    (:8-38), statement by statement.  Nested lists for the S x S tables.
    Returns (time, occupation, electrode_occupation, site_energies, traffic,
    average_occupation)."""
    n = n_sites + n_el
    occ = [float(x) for x in occupation]
    se = [float(x) for x in site_energies]
    eo = [0.0] * n_el
    traffic = [[0.0] * n for _ in range(n)]
    avg = [0.0] * n_sites
    time = 0.0
    for _ in range(hops):
        difference = [0.0] * n_sites
        for i in range(n_sites):
            ai = 0.0
            for j in range(n_sites):
                if j != i:
                    ai += (1 - occ[j]) / distances[i][j]
            se[i] = E_constant[i] - I_0 * R * ai
        trans = [[0.0] * n for _ in range(n)]
        tot = 0.0
        for i in range(n):
            for j in range(n):
                bp = prob_possible(i, j, n_sites, occ)
                if i < n_sites and j < n_sites:
                    dE = se[j] - se[i] - _fdiv(I_0 * R, distances[i][j])
                else:
                    dE = se[j] - se[i]
                t = bp * nu * math.exp(-dE / kT) if dE > 0 else bp * nu
                t *= transitions_constant[i][j]
                trans[i][j] = t
                tot += t
                if i < n_sites:
                    difference[i] -= t
                if j < n_sites:
                    difference[j] += t
        mc = 1.0
        for i in range(n_sites):
            nv = occ[i] + difference[i] / tot
            if nv < 0:
                mc = min(mc, occ[i] / -(difference[i] / tot))
            if nv > 1:
                mc = min(mc, (1 - occ[i]) / (difference[i] / tot))
        ts = 0.98 * mc / tot
        time += ts
        if record:
            for i in range(n_sites):
                avg[i] += occ[i] * ts
        for i in range(n):
            for j in range(n):
                if i >= n_sites and j >= n_sites:
                    break
                rate = trans[i][j] * mc / tot
                if i < n_sites:
                    occ[i] -= rate
                else:
                    eo[i - n_sites] -= rate
                if j < n_sites:
                    occ[j] += rate
                else:
                    eo[j - n_sites] += rate
                if record:
                    traffic[i][j] += rate
                    traffic[j][i] -= rate
        occ = [0.0 if x < 0 else (1.0 if x > 1 else x) for x in occ]
    return time, occ, eo, se, traffic, avg
