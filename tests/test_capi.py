"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT
from fir_amd import lib as firlib


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(fir_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "fir_forward" in names and "fir_osem_subiter" in names
    assert names == set(firlib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(firlib.LIB_PATH):
        pytest.fail("libfir_hip.so is not built (python -m fir_amd.build)")
    lib = ctypes.CDLL(firlib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.fir_abi_version.restype = ctypes.c_int
    assert lib.fir_abi_version() == 1


def test_load_configures_signatures():
    lib = firlib.load()
    assert lib.fir_forward.restype is ctypes.c_int
    # error reporting works without a device
    lib.fir_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.fir_last_error(), bytes)


def test_null_context_is_an_error_not_a_crash():
    lib = firlib.load()
    rc = lib.fir_set_walk_mode(None, 0)
    assert rc == 1
    assert b"NULL" in lib.fir_last_error()
