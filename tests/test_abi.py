"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls; runs without a GPU)."""
import os
import re

import pytest

from visrtx_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header: str, prefix: str):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text))
    return {n for n in names if not n.startswith("vrx_status")}


def test_hip_library_exports_every_declared_symbol():
    _native.hip()
    syms = _native.exported_symbols(_native.HIP_LIB_PATH)
    want = declared("vrx_hip.h", "vrx_")
    assert want, "no declarations parsed"
    missing = sorted(want - syms)
    assert not missing, missing


def test_anari_library_exports_every_declared_symbol():
    _native.anari_lib()
    syms = _native.exported_symbols(_native.ANARI_LIB_PATH)
    want = declared("anari/anari.h", "anari") | declared("anari/ext/visrtx_amd/visrtx_amd.h", "")
    want = {w for w in want if w.startswith(("anari", "make", "visrtx"))}
    missing = sorted(want - syms)
    assert not missing, missing


def test_every_hip_entry_point_has_64bit_safe_argtypes():
    """Each exported vrx_ function is bound with explicit argtypes, so a device
    address can never be passed as a 32-bit C int."""
    L = _native.hip()
    want = declared("vrx_hip.h", "vrx_")
    missing = sorted(n for n in want if n not in L._vrx_sigs)
    assert not missing, missing
    for name in want:
        assert getattr(L, name).argtypes is not None, name


def test_ctypes_layouts_match_the_header(tmp_path):
    """Every ctypes mirror of a vrx_hip.h struct has the C layout: a small C
    program built with gcc from the header prints sizeof and each field's
    offsetof, and they must equal the ctypes sizes and offsets (a field added
    to or removed from one side only would shift every later field)."""
    import ctypes as C
    import shutil
    import subprocess
    from visrtx_amd import _native as N
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    structs = [n for n in dir(N) if n.startswith("vrx_")
               and isinstance(getattr(N, n), type) and issubclass(getattr(N, n), C.Structure)]
    assert {"vrx_renderer", "vrx_scene", "vrx_volume", "vrx_batch", "vrx_diffvol",
            "vrx_camera"} <= set(structs)
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vrx_hip.h"', "int main(void) {"]
    for sname in structs:
        lines.append(f'  printf("{sname} size %zu\\n", sizeof({sname}));')
        for fname, _ in getattr(N, sname)._fields_:
            lines.append(f'  printf("{sname} {fname} %zu\\n", offsetof({sname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([gcc, "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        parts = ln.split()
        got[tuple(parts[:-1])] = int(parts[-1])
    for sname in structs:
        st = getattr(N, sname)
        assert got[(sname, "size")] == C.sizeof(st), sname
        for fname, _ in st._fields_:
            assert got[(sname, fname)] == getattr(st, fname).offset, (sname, fname)
