"""The C-ABI library loads and exports every symbol include/ocmp.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from opencmp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, 'include', 'ocmp.h')) as f:
        txt = f.read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(ocmp_[a-z0-9_]+)\s*\(', txt)))


def test_header_and_bindings_agree():
    decl = declared_symbols()
    assert len(decl) >= 25
    assert sorted(_lib.SIGNATURES) == decl


def test_library_loads_and_exports_everything():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail('libocmp_mi355x.so not built (run python -m opencmp_amd.build)')
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    L = _lib.load()
    assert L.ocmp_version() == 1


def test_struct_layouts_match_header():
    # sizes of the ABI structs as laid out by the C compiler (x86-64 SysV)
    assert ctypes.sizeof(_lib.FeDesc) == 24
    assert ctypes.sizeof(_lib.Field) == 24
    assert ctypes.sizeof(_lib.SolveOpts) == 48
    assert ctypes.sizeof(_lib.SolveStats) == 40
    assert ctypes.sizeof(_lib.Term) == 104
    assert ctypes.sizeof(_lib.LTerm) == 160
