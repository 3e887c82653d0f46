"""ctypes bindings of the two native libraries of the MI355X ANARI device.

``libvisrtx_amd_hip.so`` holds the gfx950 HIP kernels behind the C ABI of
``include/vrx_hip.h``; ``libanari_library_visrtx_amd.so`` is the C++ ANARI
device built on top of it (``include/anari/anari.h``).  Both are built in-tree
by ``visrtx_amd/Makefile``.  Loading fails loudly when they are missing: there
is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libvisrtx_amd_hip.so")
ANARI_LIB_PATH = os.path.join(_HERE, "libanari_library_visrtx_amd.so")


class VrxError(RuntimeError):
    pass


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise VrxError(
            f"native library {os.path.basename(path)} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or make -C visrtx_amd)")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


# ---------------------------------------------------------------------------
# C structs of include/vrx_hip.h
# ---------------------------------------------------------------------------
VRX_TRIANGLE, VRX_SPHERE, VRX_CYLINDER, VRX_BOX = 0, 1, 2, 3
VRX_H2D, VRX_D2H, VRX_D2D = 0, 1, 2
VRX_RENDER_RAYCAST, VRX_RENDER_AO = 0, 1
VRX_COLOR_F32, VRX_COLOR_SRGB8, VRX_COLOR_UNORM8 = 0, 1, 2

f3 = C.c_float * 3
f4 = C.c_float * 4
f12 = C.c_float * 12


class vrx_prims(C.Structure):
    _fields_ = [("type", C.c_int32), ("n", C.c_uint32), ("d_position", C.c_void_p),
                ("d_index", C.c_void_p), ("d_radius", C.c_void_p), ("global_radius", C.c_float)]


class vrx_bvh(C.Structure):
    _fields_ = [("d_nodes", C.c_void_p), ("d_leaves", C.c_void_p), ("d_leaf_ids", C.c_void_p),
                ("d_header", C.c_void_p)]


class vrx_bvh_sizes(C.Structure):
    _fields_ = [("nodes", C.c_size_t), ("leaves", C.c_size_t), ("leaf_ids", C.c_size_t),
                ("header", C.c_size_t), ("workspace", C.c_size_t)]


class vrx_bvh_header(C.Structure):
    _fields_ = [("lo", f3), ("hi", f3), ("height", C.c_uint32), ("root", C.c_uint32),
                ("n", C.c_uint32), ("type", C.c_int32), ("pad", C.c_uint32 * 6)]


class vrx_camera(C.Structure):
    _fields_ = [("pos", f3), ("d00", f3), ("du", f3), ("dv", f3), ("region", f4)]


class vrx_diffvol(C.Structure):
    _fields_ = [("d_density", C.c_void_p), ("dims", C.c_uint32 * 3), ("origin", f3),
                ("spacing", f3), ("d_lut", C.c_void_p), ("value_range", C.c_float * 2),
                ("unit_distance", C.c_float), ("step", C.c_float), ("early_stop", C.c_int32),
                ("checkpoint", C.c_int32), ("max_steps", C.c_uint32)]


_hip = None
_anari = None


def hip() -> C.CDLL:
    """The HIP kernel library (loaded once)."""
    global _hip
    if _hip is None:
        L = _load(HIP_LIB_PATH)
        L.vrx_last_error.restype = C.c_char_p
        for name in ["vrx_device_count", "vrx_set_device", "vrx_malloc", "vrx_free",
                     "vrx_memcpy", "vrx_memcpy_sync", "vrx_fill_u32", "vrx_stream_create",
                     "vrx_stream_sync", "vrx_device_sync", "vrx_bvh_query_sizes", "vrx_bvh_build",
                     "vrx_bvh_build_debug", "vrx_camera_perspective", "vrx_diffvol_forward",
                     "vrx_diffvol_backward", "vrx_make_rays", "vrx_render", "vrx_frame_clear",
                     "vrx_tlas_build", "vrx_tlas_query_sizes", "vrx_render_batch"]:
            getattr(L, name).restype = C.c_int
        L.vrx_render_scratch_bytes.restype = C.c_size_t
        L.vrx_diffvol_ckpt_floats.restype = C.c_size_t
        L.vrx_batch_scratch_bytes.restype = C.c_size_t
        L.vrx_gen_value_noise3.restype = C.c_float
        L.vrx_gen_value_noise3.argtypes = [C.c_uint64, C.c_float, C.c_float, C.c_float]
        L.vrx_gen_spheres.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.c_float, C.c_void_p,
                                      C.c_void_p]
        L.vrx_gen_heightfield.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                                          C.c_void_p, C.c_void_p]
        L.vrx_gen_noise_volume.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_int, C.c_float, C.c_void_p]
        L.vrx_gen_sphere_mesh_counts.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p]
        L.vrx_gen_sphere_mesh.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.c_void_p,
                                          C.c_void_p]
        _hip = L
    return _hip


def anari_lib() -> C.CDLL:
    """The ANARI device library (loaded once, after the HIP library)."""
    global _anari
    if _anari is None:
        hip()
        _anari = _load(ANARI_LIB_PATH)
    return _anari


def check(status: int, what: str = "vrx call") -> None:
    if status != 0:
        msg = hip().vrx_last_error().decode(errors="replace")
        raise VrxError(f"{what} failed ({status}): {msg}")


def exported_symbols(path: str) -> set:
    """Dynamic symbols a shared library defines (for the ABI tests)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}
