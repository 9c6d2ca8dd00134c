"""The C-ABI library loads and exports every function include/bh_gpu.h
declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from bhgpu import lib as bl

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bh_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(bh_[a-z_0-9]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    assert "bh_step" in names and "bh_build_tree" in names and "bh_accel" in names
    assert set(names) == set(bl.SIGNATURES), "bhgpu.lib.SIGNATURES must mirror the header"


def test_library_exports_every_symbol():
    if not os.path.exists(bl.LIB_PATH):
        pytest.fail("libbhgpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(bl.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.bh_max_particles() == (1 << 30) - 1


def test_records_match_header():
    from bhgpu import layout

    assert layout.PARTICLE.itemsize == 64 and layout.PARTICLE.fields["pos"][1] == 16
    assert layout.WALK_NODE.itemsize == 56 and layout.WALK_NODE.fields["first_child"][1] == 40
    assert layout.FORCE_PARAMS.itemsize == 32
