"""The ctypes mirrors in kmc_dn_amd/_lib.py have the layout of the C structs in
include/kmc_dn_amd.h: same fields in the same order, same offsets, same size.

A C program built with gcc from the header prints sizeof and offsetof of every
field; the ctypes classes must agree.  A field added to one side only (as
kmc_run_opts.state_store was, at the end of the struct) fails here on the CPU
instead of passing garbage options to the library.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from kmc_dn_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kmc_dn_amd.h")

STRUCTS = {"kmc_traj": _lib.kmc_traj, "kmc_run_opts": _lib.kmc_run_opts,
           "kmc_bucket": _lib.kmc_bucket, "kmc_GoSlice": _lib.GoSlice}


def header_fields(name):
    """Field names of `typedef struct name { ... } name;` in declaration order."""
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    m = re.search(r"typedef\s+struct\s+%s\s*\{(.*?)\}\s*%s\s*;" % (name, name), src, flags=re.S)
    assert m, name
    return [re.match(r".*?(\w+)\s*(\[[^\]]*\])?$", d.strip()).group(1)
            for d in m.group(1).split(";") if d.strip()]


def test_field_names_and_order():
    for cname, cls in STRUCTS.items():
        assert [f[0] for f in cls._fields_] == header_fields(cname), cname


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_offsets_and_sizes(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "kmc_dn_amd.h"',
             "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.dirname(HEADER),
                    str(src), "-o", str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}
    for cname, cls in STRUCTS.items():
        assert got[(cname, "sizeof")] == C.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])
