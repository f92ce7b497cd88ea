"""CPU checks of the C ABI: the library loads and exports every entry point
that include/*.h declares; the Python mirror binds every one of them.
No compute call is made (there is no GPU here)."""
import ctypes
import glob
import os
import re

import pytest

import bosss_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*int\s+(bh_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(B.LIB_PATH):
        pytest.fail("libbosss_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(B.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    names = declared_symbols()
    assert names <= set(B.SIGNATURES), sorted(names - set(B.SIGNATURES))
    B.lib()  # binds all signatures


def test_last_error_without_gpu_call():
    buf = ctypes.create_string_buffer(64)
    assert B.lib().bh_last_error(buf, 64) >= 0
