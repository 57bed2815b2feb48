"""IV sweeps (kmc_dopant_networks_utils.py:639-693): one launch per sweep."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, iv
from kmc_dn_amd import tables as T
from kmc_dn_amd.kmc import kmc_dn

pytestmark = pytest.mark.gpu


def test_iv_fp32_matches_fp64(gpu_engine):
    rng = np.random.RandomState(7)
    el = T.perimeter_electrodes(1, 1, rng.uniform(-2, 2, 8))
    m = kmc_dn(30, 3, 1, 1, 0, I_0=20, a=0.3, electrodes=el, rng=rng)
    volts = np.linspace(-4, 4, 9)
    occ0 = m.occupation.copy()
    c32, s32 = iv.IV(m, 0, volts, hops=20000, prehops=2000, replicas=64, seed0=1,
                     return_stderr=True)
    m.occupation[...] = occ0
    c64, s64 = iv.IV(m, 0, volts, hops=20000, prehops=2000, replicas=64, seed0=1,
                     precision=_lib.FP64_DENSE, return_stderr=True)
    assert c32.shape == (8, 9)
    z = np.abs(c32 - c64) / np.sqrt(s32 ** 2 + s64 ** 2 + 1e-300)
    print(f"[gpu] IV fp32 vs fp64: max |z| {z.max():.2f}", flush=True)
    assert np.all(z < 4.5)
    # raising electrode 0's potential pushes holes out of it: its arrival count falls
    assert np.corrcoef(volts, c64[0])[0, 1] < 0
    assert m.electrodes[0, 3] == volts[-1]
