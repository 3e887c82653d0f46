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
    # torch ships its own libamdhip64 (same soname, libamdhip64.so.7) and loads
    # it by a different file name; importing torch first makes the dynamic
    # linker bind this library to that one runtime, so torch's streams and
    # allocations and ours live in a single HIP runtime.  Loaded the other way
    # round, the process would hold two runtimes and a torch stream handle
    # passed to vrx_* would be meaningless.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


# ---------------------------------------------------------------------------
# C structs of include/vrx_hip.h
# ---------------------------------------------------------------------------
VRX_TRIANGLE, VRX_SPHERE, VRX_CYLINDER, VRX_BOX, VRX_CONE = 0, 1, 2, 3, 4
VRX_H2D, VRX_D2H, VRX_D2D = 0, 1, 2
VRX_RENDER_RAYCAST, VRX_RENDER_AO, VRX_RENDER_DIRECT = 0, 1, 2
VRX_LIGHT_DIRECTIONAL, VRX_LIGHT_POINT, VRX_LIGHT_SPOT = 0, 1, 2
VRX_COLOR_F32, VRX_COLOR_SRGB8, VRX_COLOR_UNORM8 = 0, 1, 2

f3 = C.c_float * 3
f4 = C.c_float * 4
f12 = C.c_float * 12


class vrx_prims(C.Structure):
    _fields_ = [("type", C.c_int32), ("n", C.c_uint32), ("d_position", C.c_void_p),
                ("d_index", C.c_void_p), ("d_radius", C.c_void_p), ("global_radius", C.c_float)]


class vrx_bvh(C.Structure):
    _fields_ = [("d_nodes", C.c_void_p), ("d_leaves", C.c_void_p), ("d_leaf_ids", C.c_void_p),
                ("d_header", C.c_void_p), ("d_wide", C.c_void_p)]


class vrx_bvh_sizes(C.Structure):
    _fields_ = [("nodes", C.c_size_t), ("leaves", C.c_size_t), ("leaf_ids", C.c_size_t),
                ("header", C.c_size_t), ("workspace", C.c_size_t)]


class vrx_bvh_header(C.Structure):
    _fields_ = [("lo", f3), ("hi", f3), ("height", C.c_uint32), ("root", C.c_uint32),
                ("n", C.c_uint32), ("type", C.c_int32), ("pad", C.c_uint32 * 6)]


class vrx_camera(C.Structure):
    _fields_ = [("pos", f3), ("d00", f3), ("du", f3), ("dv", f3), ("region", f4)]


class vrx_scene(C.Structure):
    _fields_ = [("d_geoms", C.c_void_p), ("num_geoms", C.c_uint32), ("d_entries", C.c_void_p),
                ("num_entries", C.c_uint32), ("tlas", vrx_bvh), ("tlas_height", C.c_uint32),
                ("max_blas_height", C.c_uint32), ("d_volumes", C.c_void_p),
                ("num_volumes", C.c_uint32), ("single_type", C.c_int32),
                ("max_volume_voxels", C.c_uint64), ("wide", C.c_uint32),
                ("num_lights", C.c_uint32), ("d_lights", C.c_void_p),
                ("num_majorant_grids", C.c_uint32), ("pad", C.c_uint32)]


class vrx_volume(C.Structure):
    _fields_ = [("d_data", C.c_void_p), ("dims", C.c_uint32 * 3), ("origin", f3),
                ("spacing", f3), ("d_lut", C.c_void_p), ("value_range", C.c_float * 2),
                ("unit_distance", C.c_float), ("id", C.c_uint32), ("inst_id", C.c_uint32),
                ("pad", C.c_uint32), ("d_majorant", C.c_void_p), ("d_ypairs", C.c_void_p)]


class vrx_light(C.Structure):
    _fields_ = [("type", C.c_int32), ("radiance", f3), ("pos", f3), ("dir", f3),
                ("cos_outer", C.c_float), ("cos_inner", C.c_float)]


class vrx_renderer(C.Structure):
    _fields_ = [("kind", C.c_int32), ("background", f4), ("ambient_color", f3),
                ("ambient_radiance", C.c_float), ("pixel_samples", C.c_int32),
                ("ambient_samples", C.c_int32), ("ao_distance", C.c_float),
                ("cull_backfaces", C.c_int32), ("inv_volume_sampling_rate", C.c_float),
                ("stack_lds", C.c_int32), ("vol_unroll", C.c_int32),
                ("node_width", C.c_int32), ("skip_empty", C.c_int32),
                ("volume_jitter", C.c_int32), ("defer_steps", C.c_int32)]


class vrx_batch(C.Structure):
    _fields_ = [("num_views", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("d_cams", C.c_void_p), ("d_depth", C.c_void_p), ("d_prim_id", C.c_void_p),
                ("d_obj_id", C.c_void_p), ("d_inst_id", C.c_void_p), ("d_color_f32", C.c_void_p),
                ("d_color_u32", C.c_void_p), ("d_scratch", C.c_void_p),
                ("scratch_bytes", C.c_size_t), ("d_work_counter", C.c_void_p),
                ("tile_w", C.c_int32), ("tile_h", C.c_int32), ("block_threads", C.c_int32),
                ("persistent_blocks", C.c_int32)]


VRX_ADJOINT_SUFFIX, VRX_ADJOINT_REPLAY = 0, 1


class vrx_diffvol(C.Structure):
    _fields_ = [("d_density", C.c_void_p), ("dims", C.c_uint32 * 3), ("origin", f3),
                ("spacing", f3), ("d_lut", C.c_void_p), ("value_range", C.c_float * 2),
                ("unit_distance", C.c_float), ("step", C.c_float), ("early_stop", C.c_int32),
                ("checkpoint", C.c_int32), ("max_steps", C.c_uint32),
                ("image_w", C.c_uint32), ("image_h", C.c_uint32), ("adjoint", C.c_int32)]


_hip = None
_anari = None


def hip() -> C.CDLL:
    """The HIP kernel library (loaded once)."""
    global _hip
    if _hip is None:
        L = _load(HIP_LIB_PATH)
        vp, sz, u32, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int
        # every entry point gets explicit argtypes: device addresses are 64-bit
        # and must never be passed as a default (32-bit) C int
        sigs = {
            "vrx_last_error": (C.c_char_p, []),
            "vrx_version": (i32, []),
            "vrx_device_count": (i32, [vp]),
            "vrx_set_device": (i32, [i32]),
            "vrx_get_device": (i32, [vp]),
            "vrx_device_info": (i32, [i32, vp, sz, vp, vp]),
            "vrx_malloc": (i32, [vp, sz]),
            "vrx_free": (i32, [vp]),
            "vrx_host_malloc": (i32, [vp, sz]),
            "vrx_host_free": (i32, [vp]),
            "vrx_host_device_pointer": (i32, [vp, vp]),
            "vrx_memcpy": (i32, [vp, vp, sz, i32, vp]),
            "vrx_memcpy_sync": (i32, [vp, vp, sz, i32]),
            "vrx_pointer_kind": (i32, [vp, C.POINTER(C.c_int)]),
            "vrx_fill_u32": (i32, [vp, u32, sz, vp]),
            "vrx_copy_kernel": (i32, [vp, vp, sz, vp]),
            "vrx_stream_create": (i32, [vp]),
            "vrx_stream_destroy": (i32, [vp]),
            "vrx_stream_sync": (i32, [vp]),
            "vrx_device_sync": (i32, []),
            "vrx_event_create": (i32, [vp]),
            "vrx_event_destroy": (i32, [vp]),
            "vrx_event_record": (i32, [vp, vp]),
            "vrx_event_sync": (i32, [vp]),
            "vrx_event_query": (i32, [vp]),
            "vrx_event_elapsed_ms": (i32, [vp, vp, vp]),
            "vrx_launch_host_func": (i32, [vp, vp, vp]),
            "vrx_bvh_query_sizes": (i32, [i32, u32, vp]),
            "vrx_bvh_build": (i32, [vp, vp, vp, sz, vp]),
            "vrx_bvh_build_debug": (i32, [vp, vp, vp, sz, vp, vp, vp]),
            "vrx_bvh_wide_bytes": (sz, [u32]),
            "vrx_bvh_widen": (i32, [vp, u32, vp]),
            "vrx_render_scratch_bytes": (sz, [vp, vp, u32, u32]),
            "vrx_frame_clear": (i32, [vp, vp]),
            "vrx_render": (i32, [vp, vp, vp, vp, vp]),
            "vrx_frame_collect_counters": (i32, [vp, vp, vp]),
            "vrx_camera_perspective": (i32, [vp, vp, vp, C.c_float, C.c_float, vp, vp]),
            "vrx_volume_grid_cells": (sz, [vp]),
            "vrx_volume_majorant_grid": (i32, [vp, vp, vp]),
            "vrx_volume_build_ypairs": (i32, [vp, vp, vp]),
            "vrx_volume_sample_distance": (i32, [vp, u32, vp, vp, vp, C.c_uint64, vp, vp, vp, vp]),
            "vrx_tlas_query_sizes": (i32, [u32, vp]),
            "vrx_tlas_build": (i32, [vp, vp, vp, sz, vp]),
            "vrx_diffvol_forward": (i32, [vp, u32, vp, vp, vp, vp, vp]),
            "vrx_diffvol_backward": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp]),
            "vrx_diffvol_ckpt_floats": (sz, [vp, u32]),
            "vrx_make_rays": (i32, [vp, u32, u32, u32, vp, vp, vp]),
            "vrx_batch_scratch_bytes": (sz, [vp, vp, vp]),
            "vrx_render_batch": (i32, [vp, vp, vp, vp]),
            "vrx_render_batch_static": (i32, [vp, vp, vp, vp]),
            "vrx_batch_graph_create": (i32, [vp, vp, vp, vp, vp]),
            "vrx_graph_launch": (i32, [vp, vp]),
            "vrx_graph_destroy": (i32, [vp]),
            "vrx_batch_autotune": (i32, [vp, vp, vp, vp, vp]),
            "vrx_gen_pcg32": (u32, [vp, C.c_uint64]),
            "vrx_gen_value_noise3": (C.c_float, [C.c_uint64, C.c_float, C.c_float, C.c_float]),
            "vrx_gen_spheres": (None, [C.c_uint64, u32, C.c_float, C.c_float, vp, vp]),
            "vrx_gen_heightfield": (None, [C.c_uint64, u32, u32, C.c_float, vp, vp]),
            "vrx_gen_noise_volume": (None, [C.c_uint64, u32, u32, u32, i32, C.c_float, vp]),
            "vrx_gen_sphere_mesh_counts": (None, [u32, vp, vp]),
            "vrx_gen_sphere_mesh": (None, [C.c_uint64, u32, C.c_float, vp, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._vrx_sigs = sigs
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
