"""Philox4x32-10 of the oracle: known-answer vectors and rocRAND cross-check."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors, philox4x32 with 10 rounds: counter(4), key(2) -> output(4)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_random123_kat(ctr, key, want):
    assert tuple(int(x) for x in oracle.philox4x32_10(ctr, key)) == want


def test_against_rocrand(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("hipcc not available")
    exe = tmp_path / "dump"
    subprocess.run([hipcc, "-O2", os.path.join(HERE, "golden", "rocrand_philox_dump.cpp"), "-o",
                    str(exe)], check=True, capture_output=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    n = 0
    for ln in lines:
        if not ln.strip():
            continue
        k, o, *w = (int(x) for x in ln.split())
        got = oracle.philox4x32_10([o & 0xffffffff, o >> 32, 0, 0], [k & 0xffffffff, k >> 32])
        assert [int(x) for x in got] == w, (k, o)
        n += 1
    assert n == 25


def test_stream_uniforms_layout():
    # uniforms are 53-bit doubles in [0, 1) built from words (0,1) and (2,3)
    u = oracle.stream_uniforms(7, 0, 2000)
    assert u.shape == (2000, 2)
    assert (u >= 0).all() and (u < 1).all()
    w = oracle.philox4x32_10([5, 0, 0, 0], [7, 0])
    a = ((int(w[0]) >> 5) * 67108864 + (int(w[1]) >> 6)) / 2.0 ** 53
    b = ((int(w[2]) >> 5) * 67108864 + (int(w[3]) >> 6)) / 2.0 ** 53
    assert u[5, 0] == a and u[5, 1] == b
    # moments of a uniform
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
