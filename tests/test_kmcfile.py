"""The .kmc reader decodes the pickle opcode stream as data and never unpickles."""
import pickle

import numpy as np
import pytest

from kmc_dn_amd.kmcfile import KmcFormatError, read_kmc, write_kmc


def test_round_trip(tmp_path):
    attrs = dict(N=30, R=0.18257418583505536, acceptors=np.random.RandomState(0).rand(30, 3),
                 occupation=np.array([True, False, True]), electrode_occupation=np.arange(8),
                 mean_currents=[np.float64(1.5), np.float64(-2.0)], res=0.01, name_tuple=(1, 2.0),
                 fortran=np.asfortranarray(np.arange(12.0).reshape(3, 4)), empty=np.zeros((0, 4)),
                 f32=np.arange(4, dtype=np.float32), skip_me="strings are not saved")
    p = tmp_path / "x.kmc"
    write_kmc(p, attrs)
    d = read_kmc(str(p))
    assert "skip_me" not in d
    for k, v in attrs.items():
        if k == "skip_me":
            continue
        if isinstance(v, np.ndarray):
            assert np.array_equal(d[k], v) and d[k].dtype == v.dtype and d[k].shape == v.shape, k
        else:
            assert d[k] == v, k


class _Evil:
    def __reduce__(self):
        return (print, ("pwned",))


def test_rejects_foreign_globals():
    data = pickle.dumps({"x": _Evil()}, protocol=3)
    with pytest.raises(KmcFormatError):
        read_kmc(data)


def test_rejects_object_arrays():
    data = pickle.dumps({"x": np.array([object(), 1], dtype=object)}, protocol=3)
    with pytest.raises(KmcFormatError):
        read_kmc(data)
