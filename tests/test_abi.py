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
