"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every function include/tahoe_gpu.h declares (no GPU work is done here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tahoe_gpu.h")


def declared():
    txt = open(HEADER).read()
    return re.findall(r"^\s*(?:int|const char\*)\s+(tg_\w+)\s*\(", txt, flags=re.M)


def test_header_lists_entry_points():
    names = declared()
    assert len(names) >= 25
    from tahoe_amd import api
    assert names == api.SYMBOLS


def test_library_exports_every_declared_symbol():
    from tahoe_amd import api
    if not os.path.exists(api.LIB_PATH):
        pytest.fail("libtahoe_gpu.so is not built (python -m tahoe_amd.build)")
    lib = ctypes.CDLL(api.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_library_has_gfx950_code_object():
    from tahoe_amd import api
    data = open(api.LIB_PATH, "rb").read()
    assert b"gfx950" in data
