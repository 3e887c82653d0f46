"""ANARI introspection (anariGetObjectSubtypes / anariGetObjectInfo /
anariGetParameterInfo) on the CPU: the device lists its subtypes, each
object's parameters, and the defaults and clamps it applies at commit
(renderer defaults Renderer.cpp:152-170, sphere radius Sphere.cu:52,
camera fovy 60 degrees and direction +z Camera.cpp:68-76)."""
import ctypes as C
import math

import pytest

from visrtx_amd import _native as N
from visrtx_amd import anari as A

ANARI_STRING = 101
ANARI_BOOL = 103
ANARI_PARAMETER_LIST = 152
ANARI_STRING_LIST_T = 150


class ANARIParameter(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int)]


@pytest.fixture(scope="module")
def lib():
    L = N.anari_lib()
    L.anariGetObjectInfo.restype = C.c_void_p
    L.anariGetObjectInfo.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    L.anariGetParameterInfo.restype = C.c_void_p
    L.anariGetParameterInfo.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int,
                                        C.c_char_p, C.c_int]
    L.anariGetObjectSubtypes.restype = C.POINTER(C.c_char_p)
    L.anariGetObjectSubtypes.argtypes = [C.c_void_p, C.c_int]
    return L


def params(L, typ, sub):
    p = L.anariGetObjectInfo(None, typ, sub, b"parameter", ANARI_PARAMETER_LIST)
    assert p
    arr = C.cast(p, C.POINTER(ANARIParameter))
    out = {}
    i = 0
    while arr[i].name:
        out[arr[i].name.decode()] = arr[i].type
        i += 1
    return out


def default(L, typ, sub, name, ptype, ctype, n=1):
    p = L.anariGetParameterInfo(None, typ, sub, name.encode(), ptype, b"default", ptype)
    if not p:
        return None
    v = C.cast(p, C.POINTER(ctype * n)).contents
    return list(v) if n > 1 else v[0]


def subtypes(L, typ):
    p = L.anariGetObjectSubtypes(None, typ)
    out, i = [], 0
    while p[i]:
        out.append(p[i].decode())
        i += 1
    return out


def test_subtypes(lib):
    assert set(subtypes(lib, A.GEOMETRY)) >= {"triangle", "quad", "sphere", "cylinder", "cone", "curve"}
    assert set(subtypes(lib, A.RENDERER)) >= {"raycast", "ao", "directLight", "default"}
    assert set(subtypes(lib, A.LIGHT)) == {"directional", "point", "spot"}


def test_every_listed_subtype_has_a_parameter_table(lib):
    for typ in (A.GEOMETRY, A.RENDERER, A.LIGHT):
        for sub in subtypes(lib, typ):
            assert params(lib, typ, sub.encode()), (typ, sub)


def test_parameter_lists(lib):
    sph = params(lib, A.GEOMETRY, b"sphere")
    assert sph["vertex.position"] == A.ARRAY1D and sph["radius"] == A.FLOAT32
    ren = params(lib, A.RENDERER, b"ao")
    for k in ("background", "ambientSamples", "ambientOcclusionDistance", "volumeSamplingRate"):
        assert k in ren
    assert "ambientSamples" not in params(lib, A.RENDERER, b"raycast")
    assert params(lib, A.SURFACE, None)["geometry"] == A.GEOMETRY
    req = lib.anariGetParameterInfo(None, A.GEOMETRY, b"sphere", b"vertex.position", A.ARRAY1D,
                                    b"required", ANARI_BOOL)
    assert C.cast(req, C.POINTER(C.c_int32))[0] == 1
    desc = lib.anariGetObjectInfo(None, A.VOLUME, b"transferFunction1D", b"description",
                                  ANARI_STRING)
    assert C.cast(desc, C.c_char_p).value


def test_defaults_and_clamps(lib):
    assert default(lib, A.GEOMETRY, b"sphere", "radius", A.FLOAT32, C.c_float) == \
        pytest.approx(0.01)
    assert default(lib, A.GEOMETRY, b"cylinder", "radius", A.FLOAT32, C.c_float) == 1.0
    assert default(lib, A.RENDERER, b"raycast", "volumeSamplingRate", A.FLOAT32, C.c_float) == 0.125
    lo = lib.anariGetParameterInfo(None, A.RENDERER, b"raycast", b"volumeSamplingRate", A.FLOAT32,
                                   b"minimum", A.FLOAT32)
    hi = lib.anariGetParameterInfo(None, A.RENDERER, b"raycast", b"volumeSamplingRate", A.FLOAT32,
                                   b"maximum", A.FLOAT32)
    assert C.cast(lo, C.POINTER(C.c_float))[0] == pytest.approx(1e-3)
    assert C.cast(hi, C.POINTER(C.c_float))[0] == 10.0
    assert default(lib, A.RENDERER, b"ao", "ambientSamples", A.INT32, C.c_int32) == 1
    assert default(lib, A.RENDERER, b"ao", "background", A.FLOAT32_VEC4, C.c_float, 4) == \
        [0.0, 0.0, 0.0, 1.0]
    assert default(lib, A.CAMERA, b"perspective", "direction", A.FLOAT32_VEC3, C.c_float, 3) == \
        [0.0, 0.0, 1.0]
    assert default(lib, A.CAMERA, b"perspective", "fovy", A.FLOAT32, C.c_float) == \
        pytest.approx(1.0471976)
    assert default(lib, A.SURFACE, None, "id", A.UINT32, C.c_uint32) == 0xFFFFFFFF
    # ambientRadiance defaults: 1 for ao, 0 for directLight / default
    # (AmbientOcclusion.cpp:44, DirectLight.cpp:46)
    assert default(lib, A.RENDERER, b"ao", "ambientRadiance", A.FLOAT32, C.c_float) == 1.0
    for sub in (b"directLight", b"default"):
        assert default(lib, A.RENDERER, sub, "ambientRadiance", A.FLOAT32, C.c_float) == 0.0
        assert default(lib, A.RENDERER, sub, "lightFalloff", A.FLOAT32, C.c_float) == 1.0
    assert default(lib, A.LIGHT, b"spot", "openingAngle", A.FLOAT32, C.c_float) == \
        pytest.approx(math.pi)
    assert default(lib, A.LIGHT, b"spot", "falloffAngle", A.FLOAT32, C.c_float) == \
        pytest.approx(0.1)
    assert default(lib, A.LIGHT, b"directional", "direction", A.FLOAT32_VEC3, C.c_float, 3) == \
        [0.0, 0.0, -1.0]
    assert params(lib, A.WORLD, None)["light"] == A.ARRAY1D
    for name in (b"transform", b"id"):   # transform / id arrays (Instance.cpp:115-128)
        assert lib.anariGetParameterInfo(None, A.INSTANCE, b"transform", name, A.ARRAY1D,
                                         b"description", ANARI_STRING)
    assert params(lib, A.GROUP, None)["light"] == A.ARRAY1D
    # wrong parameter type or unknown names answer null
    assert not lib.anariGetParameterInfo(None, A.GEOMETRY, b"sphere", b"radius", A.INT32,
                                         b"default", A.INT32)
    assert not lib.anariGetObjectInfo(None, A.GEOMETRY, b"isosurface", b"parameter",
                                      ANARI_PARAMETER_LIST)


# Parameters of the reference's renderers that this device does not
# implement: the OptiX denoiser and checkerboarded sampling are outside the
# ported path.
_NOT_BUILT = {"denoise", "checkerboarding"}
_TYPES = {"ANARI_INT32": (A.INT32, C.c_int32), "ANARI_FLOAT32": (A.FLOAT32, C.c_float),
          "ANARI_BOOL": (ANARI_BOOL, C.c_int32)}


def test_renderer_tables_match_the_reference_device_declaration(lib):
    """Every parameter the reference declares for the default, ao and
    raycast renderers (devices/rtx/visrtx_device.json, committed as
    tests/golden/renderer_params.json) is listed here with the same type and
    default, except the two features not built."""
    import json
    import os
    spec = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "renderer_params.json")))
    assert {"default", "ao", "raycast"} <= set(spec)
    checked = 0
    for sub in ("default", "ao", "raycast"):
        mine = params(lib, A.RENDERER, sub.encode())
        for p in spec[sub]:
            if p["name"] in _NOT_BUILT:
                continue
            ptype, ctype = _TYPES[p["type"]]
            assert mine.get(p["name"]) == ptype, (sub, p["name"])
            got = default(lib, A.RENDERER, sub.encode(), p["name"], ptype, ctype)
            if ptype == ANARI_BOOL:
                assert bool(got) == p["default"], (sub, p["name"])
            else:
                assert got == pytest.approx(p["default"], rel=1e-6), (sub, p["name"], got)
            checked += 1
    assert checked == 16


def _strings(p):
    arr = C.cast(p, C.POINTER(C.c_char_p))
    out, i = [], 0
    while arr[i]:
        out.append(arr[i].decode())
        i += 1
    return out


def test_extension_and_feature_queries(lib):
    """The device's "extension" and "subtypes.renderer" string-list
    properties (VisRTXDevice.cpp:541-546) and each object's "feature" list
    (read by visrtxGetObjectExtensions, VisRTXFeatureUtility.cpp:64-77): the
    triangle back-face culling extension the reference reports is there, the
    CUDA-only ones are not, and every object feature is a device extension."""
    dev = A.Device()
    p = C.c_void_p()
    L = dev.L
    assert L.anariGetProperty(dev.h, dev.h, b"extension", ANARI_STRING_LIST_T, C.byref(p),
                              C.sizeof(p), A.WAIT) == 1
    ext = _strings(p.value)
    assert "ANARI_VISRTX_TRIANGLE_BACK_FACE_CULLING" in ext
    assert "ANARI_KHR_GEOMETRY_SPHERE" in ext and "ANARI_KHR_ARRAY1D_REGION" in ext
    assert "ANARI_VISRTX_ARRAY_CUDA" not in ext and "ANARI_NV_ARRAY_CUDA" not in ext
    assert len(ext) == len(set(ext))
    assert L.anariGetProperty(dev.h, dev.h, b"subtypes.renderer", ANARI_STRING_LIST_T, C.byref(p),
                              C.sizeof(p), A.WAIT) == 1
    assert set(_strings(p.value)) == set(subtypes(lib, A.RENDERER))
    for typ in (A.GEOMETRY, A.RENDERER, A.LIGHT):
        for sub in subtypes(lib, typ):
            f = lib.anariGetObjectInfo(None, typ, sub.encode(), b"feature", ANARI_STRING_LIST_T)
            assert f, (typ, sub)
            feats = _strings(f)
            assert feats and set(feats) <= set(ext), (typ, sub, feats)
    tri = _strings(lib.anariGetObjectInfo(None, A.GEOMETRY, b"triangle", b"feature",
                                          ANARI_STRING_LIST_T))
    assert tri == ["ANARI_KHR_GEOMETRY_TRIANGLE", "ANARI_VISRTX_TRIANGLE_BACK_FACE_CULLING"]
    dev.close()
