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


def test_descriptor_layout_matches_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset gcc gives it in
    include/jpgpu.h."""
    import subprocess
    from j_presto_amd.abi import SysDesc, EnergyLog
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "jpgpu.h"', 'int main(void) {']
    for cname, ct in (("jp_sys_desc", SysDesc), ("jp_energy_log", EnergyLog)):
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            cf = "lambda" if f == "lambda_" else f
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {cf}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = {(a, b): int(c) for a, b, c in (ln.split() for ln in out.splitlines())}
    for cname, ct in (("jp_sys_desc", SysDesc), ("jp_energy_log", EnergyLog)):
        assert got[(cname, "sizeof")] == ctypes.sizeof(ct), cname
        for f, _ in ct._fields_:
            assert got[(cname, f)] == getattr(ct, f).offset, (cname, f)
