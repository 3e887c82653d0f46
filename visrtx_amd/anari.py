"""Minimal Python binding of the device's ANARI C API (ctypes).

It mirrors the shape of ANARI client code such as
``examples/simple/testApp_spheres.cpp:167-263`` in the VisRTX tree: create a
device, arrays, geometry, surfaces, a world, camera, renderer and frame,
render, map channels.  Arrays are created as *managed* ANARI arrays (device
owned memory filled through map/unmap), so numpy buffers need not outlive the
call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native

# ANARI data types (include/anari/anari.h)
UNKNOWN = 0
DATA_TYPE = 100
STRING = 101
VOID_POINTER = 102
BOOL = 103
FRAME_COMPLETION_CALLBACK = 203
ARRAY1D, ARRAY2D, ARRAY3D = 504, 505, 506
CAMERA, FRAME, GEOMETRY, GROUP, INSTANCE = 507, 508, 509, 510, 511
LIGHT, MATERIAL, RENDERER, SURFACE, SAMPLER = 512, 513, 514, 515, 516
SPATIAL_FIELD, VOLUME, WORLD = 517, 518, 519
INT32 = 1016
UINT32, UINT32_VEC2, UINT32_VEC3, UINT32_VEC4 = 1020, 1021, 1022, 1023
UINT64 = 1028
UFIXED8 = 1036
UFIXED8_VEC4 = 1039
FLOAT32, FLOAT32_VEC2, FLOAT32_VEC3, FLOAT32_VEC4 = 1068, 1069, 1070, 1071
UFIXED8_RGBA_SRGB = 2003
FLOAT32_BOX1, FLOAT32_BOX2, FLOAT32_BOX3 = 2004, 2005, 2006
FLOAT32_MAT4, FLOAT32_MAT3x4 = 2026, 2028
NO_WAIT, WAIT = 0, 1

_OBJ_TYPES = {ARRAY1D, ARRAY2D, ARRAY3D, CAMERA, FRAME, GEOMETRY, GROUP, INSTANCE, LIGHT,
              MATERIAL, RENDERER, SURFACE, SAMPLER, SPATIAL_FIELD, VOLUME, WORLD}

_ELEM = {  # element type -> (numpy dtype, components)
    FLOAT32: (np.float32, 1), FLOAT32_VEC2: (np.float32, 2), FLOAT32_VEC3: (np.float32, 3),
    FLOAT32_VEC4: (np.float32, 4), UINT32: (np.uint32, 1), UINT32_VEC2: (np.uint32, 2),
    UINT32_VEC3: (np.uint32, 3), UINT32_VEC4: (np.uint32, 4), UFIXED8: (np.uint8, 1),
    INT32: (np.int32, 1), FLOAT32_MAT4: (np.float32, 16),
}

STATUS_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                        C.c_char_p)
FRAME_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)


def _api():
    L = _native.anari_lib()
    if not getattr(L, "_vrx_typed", False):
        vp = C.c_void_p
        L.makeVisRTXAMDDevice.restype = vp
        L.makeVisRTXAMDDevice.argtypes = [STATUS_CB, vp]
        L.anariLoadLibrary.restype = vp
        L.anariNewDevice.restype = vp
        L.anariNewDevice.argtypes = [vp, C.c_char_p]
        for n in ["anariNewCamera", "anariNewGeometry", "anariNewSpatialField", "anariNewVolume",
                  "anariNewMaterial", "anariNewSampler", "anariNewInstance", "anariNewRenderer",
                  "anariNewLight"]:
            getattr(L, n).restype = vp
            getattr(L, n).argtypes = [vp, C.c_char_p]
        for n in ["anariNewSurface", "anariNewGroup", "anariNewWorld", "anariNewFrame"]:
            getattr(L, n).restype = vp
            getattr(L, n).argtypes = [vp]
        L.anariNewArray1D.restype = vp
        L.anariNewArray1D.argtypes = [vp, vp, vp, vp, C.c_int, C.c_uint64]
        L.anariNewArray2D.restype = vp
        L.anariNewArray2D.argtypes = [vp, vp, vp, vp, C.c_int, C.c_uint64, C.c_uint64]
        L.anariNewArray3D.restype = vp
        L.anariNewArray3D.argtypes = [vp, vp, vp, vp, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64]
        L.anariMapArray.restype = vp
        L.anariMapArray.argtypes = [vp, vp]
        L.anariUnmapArray.argtypes = [vp, vp]
        L.anariSetParameter.argtypes = [vp, vp, C.c_char_p, C.c_int, vp]
        L.anariUnsetParameter.argtypes = [vp, vp, C.c_char_p]
        L.anariCommitParameters.argtypes = [vp, vp]
        L.anariRelease.argtypes = [vp, vp]
        L.anariRetain.argtypes = [vp, vp]
        L.anariGetProperty.restype = C.c_int
        L.anariGetProperty.argtypes = [vp, vp, C.c_char_p, C.c_int, vp, C.c_uint64, C.c_int]
        L.anariMapFrame.restype = vp
        L.anariMapFrame.argtypes = [vp, vp, C.c_char_p, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        L.anariUnmapFrame.argtypes = [vp, vp, C.c_char_p]
        L.anariRenderFrame.argtypes = [vp, vp]
        L.anariFrameReady.restype = C.c_int
        L.anariFrameReady.argtypes = [vp, vp, C.c_int]
        L.anariGetObjectSubtypes.restype = C.POINTER(C.c_char_p)
        L.anariGetObjectSubtypes.argtypes = [vp, C.c_int]
        L._vrx_typed = True
    return L


class Device:
    """One ANARI device (``makeVisRTXAMDDevice``)."""

    def __init__(self, verbose: bool = False):
        self.L = _api()
        self.messages: list[tuple[int, str]] = []

        def _cb(user, dev, src, stype, severity, code, msg):
            text = msg.decode(errors="replace") if msg else ""
            self.messages.append((severity, text))
            if verbose or severity <= 2:
                print(f"[visrtx_amd:{severity}] {text}", flush=True)

        self._cb = STATUS_CB(_cb)
        self.h = self.L.makeVisRTXAMDDevice(self._cb, None)
        if not self.h:
            raise _native.VrxError("makeVisRTXAMDDevice failed")
        self._frame_cbs = []

    # -- objects -------------------------------------------------------------
    def new(self, kind: str, subtype: str | None = None):
        fn = getattr(self.L, "anariNew" + kind)
        h = fn(self.h, subtype.encode()) if subtype is not None else fn(self.h)
        if not h:
            raise _native.VrxError(f"anariNew{kind}({subtype}) failed: {self.errors()}")
        return h

    def array(self, data: np.ndarray, elem: int, dims=None):
        """Managed array filled from ``data``; elem is an ANARI element type."""
        if elem in _OBJ_TYPES:
            handles = np.asarray([int(h) for h in data], dtype=np.uint64)
            n = len(handles)
            a = self.L.anariNewArray1D(self.h, None, None, None, elem, n)
            ptr = self.L.anariMapArray(self.h, a)
            C.memmove(ptr, handles.ctypes.data, handles.nbytes)
            self.L.anariUnmapArray(self.h, a)
            return a
        dt, comps = _ELEM[elem]
        arr = np.ascontiguousarray(data, dtype=dt)
        if dims is None:
            dims = (arr.size // comps,)
        if len(dims) == 1:
            a = self.L.anariNewArray1D(self.h, None, None, None, elem, dims[0])
        elif len(dims) == 2:
            a = self.L.anariNewArray2D(self.h, None, None, None, elem, dims[0], dims[1])
        else:
            a = self.L.anariNewArray3D(self.h, None, None, None, elem, dims[0], dims[1], dims[2])
        if not a:
            raise _native.VrxError(f"anariNewArray failed: {self.errors()}")
        ptr = self.L.anariMapArray(self.h, a)
        C.memmove(ptr, arr.ctypes.data, arr.nbytes)
        self.L.anariUnmapArray(self.h, a)
        return a

    def shared_array(self, ptr: int, elem: int, n: int, deleter=None, user=None):
        """A shared (or, with a deleter, captured) 1D array over the caller's
        memory at ``ptr``: host, device or managed memory
        (array/Array.cpp:38-118).  The caller keeps the memory alive."""
        a = self.L.anariNewArray1D(self.h, C.c_void_p(ptr), deleter, user, elem, n)
        if not a:
            raise _native.VrxError(f"anariNewArray1D(shared) failed: {self.errors()}")
        return a

    def set(self, obj, name: str, value, typ: int):
        if typ == STRING:
            buf = C.c_char_p(value.encode())
            self.L.anariSetParameter(self.h, obj, name.encode(), typ, C.cast(buf, C.c_void_p))
            return
        if typ in _OBJ_TYPES:
            h = C.c_void_p(value)
            self.L.anariSetParameter(self.h, obj, name.encode(), typ, C.byref(h))
            return
        if typ == DATA_TYPE or typ == INT32 or typ == BOOL:
            v = C.c_int32(int(value))
            self.L.anariSetParameter(self.h, obj, name.encode(), typ, C.byref(v))
            return
        if typ in (UINT32,):
            v = C.c_uint32(int(value))
            self.L.anariSetParameter(self.h, obj, name.encode(), typ, C.byref(v))
            return
        if typ == UINT64:
            v = C.c_uint64(int(value))
            self.L.anariSetParameter(self.h, obj, name.encode(), typ, C.byref(v))
            return
        if typ in (FRAME_COMPLETION_CALLBACK, VOID_POINTER):
            v = C.c_void_p(value)
            self.L.anariSetParameter(self.h, obj, name.encode(), typ, C.byref(v))
            return
        dt = np.uint32 if typ in (UINT32_VEC2, UINT32_VEC3) else np.float32
        arr = np.ascontiguousarray(np.asarray(value, dtype=dt).reshape(-1))
        self.L.anariSetParameter(self.h, obj, name.encode(), typ, arr.ctypes.data)

    def commit(self, obj=None):
        self.L.anariCommitParameters(self.h, obj if obj is not None else self.h)

    def release(self, obj):
        self.L.anariRelease(self.h, obj)

    def retain(self, obj):
        self.L.anariRetain(self.h, obj)

    def get_property(self, obj, name: str, typ: int, wait: bool = True):
        if typ == FLOAT32:
            v = C.c_float()
        elif typ in (INT32, BOOL):
            v = C.c_int32()
        elif typ == UINT64:
            v = C.c_uint64()
        elif typ == FLOAT32_BOX3:
            v = (C.c_float * 6)()
        elif typ == VOID_POINTER:
            v = C.c_void_p()
        else:
            raise ValueError(typ)
        ok = self.L.anariGetProperty(self.h, obj if obj is not None else self.h, name.encode(),
                                     typ, C.byref(v), C.sizeof(v), WAIT if wait else NO_WAIT)
        if not ok:
            return None
        if typ == FLOAT32_BOX3:
            return np.array(list(v), dtype=np.float32)
        return v.value

    def subtypes(self, typ: int):
        p = self.L.anariGetObjectSubtypes(self.h, typ)
        out = []
        i = 0
        while p[i]:
            out.append(p[i].decode())
            i += 1
        return out

    # -- frames --------------------------------------------------------------
    def render(self, frame):
        self.L.anariRenderFrame(self.h, frame)

    def ready(self, frame, wait=True) -> bool:
        return bool(self.L.anariFrameReady(self.h, frame, WAIT if wait else NO_WAIT))

    def map_frame(self, frame, channel: str):
        w, h, t = C.c_uint32(), C.c_uint32(), C.c_int()
        ptr = self.L.anariMapFrame(self.h, frame, channel.encode(), C.byref(w), C.byref(h),
                                   C.byref(t))
        if not ptr:
            return None, t.value
        n = w.value * h.value
        if t.value == FLOAT32_VEC4:
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(n * 4,)).copy()
            a = a.reshape(h.value, w.value, 4)
        elif t.value == FLOAT32:
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(n,)).copy()
            a = a.reshape(h.value, w.value)
        else:
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint32)), shape=(n,)).copy()
            a = a.reshape(h.value, w.value)
        self.L.anariUnmapFrame(self.h, frame, channel.encode())
        return a, t.value

    def map_frame_device_ptr(self, frame, channel: str) -> int:
        w, h, t = C.c_uint32(), C.c_uint32(), C.c_int()
        ptr = self.L.anariMapFrame(self.h, frame, (channel + "GPU").encode(), C.byref(w),
                                   C.byref(h), C.byref(t))
        return ptr or 0

    def errors(self):
        return [m for s, m in self.messages if s <= 2]

    def close(self):
        if self.h:
            self.L.anariRelease(self.h, self.h)
            self.h = None
