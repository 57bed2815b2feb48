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
