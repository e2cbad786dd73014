"""The C ABI boundary without a GPU: the library loads and exports every function include/ptg.h
declares, and the ctypes signatures cover exactly those functions (no compute calls here)."""

import ctypes
import os
import re
import shutil
import subprocess

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
    assert ctypes.sizeof(_lib.TreeArrays) == 15 * 8
    assert ctypes.sizeof(_lib.WalkTree) == 4 * 8 + 8 + 4 + 4 + 3 * 8


@pytest.mark.skipif(not shutil.which("gcc"), reason="gcc not available")
def test_struct_field_offsets_match_the_header(tmp_path):
    """The C compiler's sizeof/offsetof of every header struct equal the ctypes mirror's."""
    from pytreegrav_amd import _lib

    mirror = {"ptg_build_info": _lib.BuildInfo, "ptg_tree_arrays": _lib.TreeArrays, "ptg_forest": _lib.Forest,
              "ptg_walk_tree": _lib.WalkTree, "ptg_criterion": _lib.Criterion, "ptg_build_counts": _lib.BuildCounts}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "ptg.h"', "int main(void) {"]
    expect = []
    for cname, cls in mirror.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(cls))
        for field, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {field}));')
            expect.append(getattr(cls, field).offset)
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect
