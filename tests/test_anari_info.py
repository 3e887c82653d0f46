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
