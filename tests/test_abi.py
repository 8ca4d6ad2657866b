"""The C-ABI library loads and exports every entry point include/*.h declares,
and the Python-side record layouts match the header (no GPU needed)."""
from __future__ import annotations

import ctypes
import glob
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from fastfermion_amd import _abi, _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ffx_\w+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 10
    lib = _native.load()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"
    assert set(_native.EXPORTED) <= set(names)


def test_record_layouts_match_header():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "fastfermion_amd.h"
    int main(void) {
        printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(ffx_pauli_gate), sizeof(ffx_majorana_gate),
               sizeof(ffx_options), sizeof(ffx_terms), sizeof(ffx_profile),
               offsetof(ffx_pauli_gate, theta), offsetof(ffx_options, mincoeff));
        return 0;
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == _abi.PAULI_GATE_DTYPE.itemsize
    assert out[1] == _abi.MAJORANA_GATE_DTYPE.itemsize
    assert out[2] == ctypes.sizeof(_native.Options)
    assert out[3] == ctypes.sizeof(_native.Terms)
    assert out[4] == ctypes.sizeof(_native.Profile)
    assert out[5] == _abi.PAULI_GATE_DTYPE.fields["theta"][1]
    assert out[6] == _native.Options.mincoeff.offset


def test_no_device_is_reported_not_faked():
    n = _native.device_count()
    if n > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError):
        _native.Context(0)
