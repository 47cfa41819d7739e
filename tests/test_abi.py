"""libjpgpu.so loads on a host without a GPU and exports every entry point
include/jpgpu.h declares (no compute calls here)."""
import ctypes
import os
import re

from j_presto_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "jpgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(jp_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(native.SIGNATURES)


def test_library_exports_every_symbol():
    lib = native.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.jp_version().startswith(b"jpgpu")


def test_descriptor_layout_matches_header():
    from j_presto_amd.abi import SysDesc, EnergyLog
    # natom..nfrg (4 ints), 8 pointers, nexcl + ptr, nexcess + ptr, nshake + 3 ptrs, ...
    assert SysDesc.celwal.offset % 8 == 0
    assert ctypes.sizeof(EnergyLog) == 8 + 3 * 8 + 6 * 8
