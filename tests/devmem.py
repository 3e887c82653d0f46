"""Device buffers through the C ABI (vrx_malloc / vrx_memcpy) for GPU tests."""
import ctypes as C

import numpy as np

from visrtx_amd import _native


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.L = _native.hip()
        self.nbytes = max(int(nbytes), 4)
        self.ptr = C.c_void_p()
        _native.check(self.L.vrx_malloc(C.byref(self.ptr), C.c_size_t(self.nbytes)), "vrx_malloc")

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _native.check(b.L.vrx_memcpy_sync(b.ptr, C.c_void_p(a.ctypes.data),
                                              C.c_size_t(a.nbytes), _native.VRX_H2D), "h2d")
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            _native.check(self.L.vrx_memcpy_sync(C.c_void_p(out.ctypes.data), self.ptr,
                                                 C.c_size_t(out.nbytes), _native.VRX_D2H), "d2h")
        return out

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def free(self):
        if self.ptr.value:
            self.L.vrx_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
