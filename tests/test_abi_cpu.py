"""The C ABI boundary without a GPU: the library loads and exports every function include/ptg.h
declares, and the ctypes signatures cover exactly those functions (no compute calls here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pytreegrav_amd", "libptg_hip.so")


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ptg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?\**\s*\b(ptg_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("ptg_build_sort", "ptg_build_tree", "ptg_walk", "ptg_direct", "ptg_morton_order",
                 "ptg_pack_walk", "ptg_target_groups", "ptg_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libptg_hip.so not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_signatures_match_the_header():
    from pytreegrav_amd import _lib

    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_struct_layouts():
    from pytreegrav_amd import _lib

    assert ctypes.sizeof(_lib.BuildInfo) == 8 + 8 + 4 * 4 + 3 * 8 + 8
    assert ctypes.sizeof(_lib.TreeArrays) == 13 * 8
    assert ctypes.sizeof(_lib.WalkTree) == 4 * 8 + 8 + 4 + 4 + 3 * 8
