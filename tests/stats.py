"""Statistics shared by the parity tests.

* :func:`compare` -- ensemble means of two sets of per-trajectory currents,
  the standard error of their difference (independent samples), |z| and the
  relative difference.
* :func:`assert_within` -- the 3-sigma bar: every resolvable current
  (|I| > 10 standard errors) within 3 sigma; the currents that are not
  resolvable are counted and also held to 3 sigma on their difference.
* :func:`e_db` (from kmc_dn_amd.validate) -- the reference's accuracy
  statistic, the mean Bhattacharyya distance between two (mean, std) current
  estimates per electrode, thesis_indrek/validate_tests.py:80-135.
"""
from __future__ import annotations

import numpy as np


def compare(ca, cb):
    """ca (Ka, P), cb (Kb, P) per-trajectory currents -> dict of per-electrode arrays."""
    ca, cb = np.asarray(ca, float), np.asarray(cb, float)
    ma, mb = ca.mean(0), cb.mean(0)
    se = np.sqrt(ca.var(0, ddof=1) / len(ca) + cb.var(0, ddof=1) / len(cb))
    diff = np.abs(ma - mb)
    z = np.where(se > 0, diff / np.where(se > 0, se, 1.0), np.where(diff > 0, np.inf, 0.0))
    rel = diff / np.maximum(np.abs(mb), 1e-300)
    resolvable = np.abs(mb) > 10 * se
    return {"a": ma, "b": mb, "se": se, "z": z, "rel": rel, "resolvable": resolvable}


def assert_within(c, label, nsig=3.0, rel_bar=None):
    """The 3-sigma bar on every current of compare() output ``c`` (arrays of
    any shape); with ``rel_bar`` also |rel| < rel_bar on the resolvable ones."""
    z, res = np.asarray(c["z"]), np.asarray(c["resolvable"])
    msg = (f"{label}: {res.sum()} resolvable of {res.size} currents, max |z| "
           f"{z.max():.2f} (resolvable {z[res].max() if res.any() else 0:.2f})")
    if res.any():
        msg += f", max rel diff (resolvable) {np.asarray(c['rel'])[res].max():.2e}"
    print("[stat] " + msg, flush=True)
    bad = z >= nsig
    assert not bad.any(), f"{msg}: |z| >= {nsig} at {np.argwhere(bad).tolist()}"
    if rel_bar is not None and res.any():
        rel = np.asarray(c["rel"])
        assert (rel[res] < rel_bar).all(), f"{msg}: relative difference above {rel_bar}"
    return msg


# the fixture harness's statistic lives with the engine's measurement tool
from kmc_dn_amd.validate import calc_d, e_db, mean_std_pop  # noqa: E402,F401
