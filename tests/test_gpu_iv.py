"""IV sweeps (kmc_dopant_networks_utils.py:639-693): one launch per sweep."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, iv
from kmc_dn_amd import tables as T
from kmc_dn_amd.kmc import kmc_dn
from tests.stats import assert_within

pytestmark = pytest.mark.gpu


def test_iv_fp32_matches_fp64(gpu_engine):
    rng = np.random.RandomState(7)
    el = T.perimeter_electrodes(1, 1, rng.uniform(-2, 2, 8))
    m = kmc_dn(30, 3, 1, 1, 0, I_0=20, a=0.3, electrodes=el, rng=rng)
    volts = np.linspace(-4, 4, 9)
    occ0 = m.occupation.copy()
    c32, s32 = iv.IV(m, 0, volts, hops=20000, prehops=2000, replicas=64, seed0=1,
                     return_stderr=True, start="model")
    m.occupation[...] = occ0
    c64, s64 = iv.IV(m, 0, volts, hops=20000, prehops=2000, replicas=64, seed0=1,
                     precision=_lib.FP64_DENSE, return_stderr=True, start="model")
    assert c32.shape == (8, 9)
    se = np.sqrt(s32 ** 2 + s64 ** 2)
    diff = np.abs(c32 - c64)
    z = np.where(se > 0, diff / np.where(se > 0, se, 1), np.where(diff > 0, np.inf, 0))
    assert_within({"z": z, "resolvable": np.abs(c64) > 10 * se, "rel": diff / np.maximum(
        np.abs(c64), 1e-300)}, "IV fp32 vs fp64 (9 voltages x 8 electrodes)")
    # raising electrode 0's potential pushes holes out of it: its arrival count falls
    assert np.corrcoef(volts, c64[0])[0, 1] < 0
    assert m.electrodes[0, 3] == volts[-1]
