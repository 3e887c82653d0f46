"""Driver of the batched multi-view entry points of ``include/vrx_hip.h``.

``vrx_render_batch`` renders many camera poses of one ANARI world in a single
persistent launch (the reference issues one optixLaunch per frame,
frame/Frame.cu:282-289, and has no batched form).  This module only wires
device memory and streams (torch) to the C ABI: the scene and renderer
descriptions come from the ANARI device's ``vrxScene`` / ``vrxRenderer``
extension properties, so the batched frames trace exactly the BVHs that the
device built for the world.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from . import anari as A


def scene_of(dev: A.Device, world, wide: bool = False) -> N.vrx_scene:
    """Copy of the world's device-side scene description (rebuilds if needed);
    ``wide`` also builds the 4-wide BVH nodes so node width 4 can be used."""
    p = dev.get_property(world, "vrxSceneWide" if wide else "vrxScene", A.VOID_POINTER)
    if not p:
        raise N.VrxError("world has no scene (vrxScene property unavailable)")
    return N.vrx_scene.from_buffer_copy(C.string_at(p, C.sizeof(N.vrx_scene)))


def renderer_of(dev: A.Device, renderer) -> N.vrx_renderer:
    p = dev.get_property(renderer, "vrxRenderer", A.VOID_POINTER)
    if not p:
        raise N.VrxError("renderer has no vrxRenderer property")
    return N.vrx_renderer.from_buffer_copy(C.string_at(p, C.sizeof(N.vrx_renderer)))


def camera_struct(cam) -> N.vrx_camera:
    """vrx_camera of a scenes.CameraDesc (camera/Perspective.cpp:40-72)."""
    c = N.vrx_camera()
    f3 = C.c_float * 3
    N.check(N.hip().vrx_camera_perspective(f3(*cam.position), f3(*cam.direction), f3(*cam.up),
                                           float(cam.fovy), float(cam.aspect), None, C.byref(c)),
            "vrx_camera_perspective")
    return c


class BatchFrame:
    """Output channels of ``views`` frames of ``width`` x ``height`` and the
    launch state of one batched render (queue counter, stack-spill scratch).

    Channel layouts match the per-view frame: element (v, y, x), row 0 at the
    bottom; color is float32 RGBA, or sRGB8 packed in uint32 when
    ``srgb=True``."""

    def __init__(self, scene: N.vrx_scene, renderer: N.vrx_renderer, cams, width: int,
                 height: int, srgb: bool = False, device="cuda"):
        self.scene = scene
        self.renderer = N.vrx_renderer.from_buffer_copy(renderer)
        if self.renderer.node_width == 4 and not scene.wide:
            self.renderer.node_width = 2   # the scene was fetched without 4-wide nodes
        structs = [camera_struct(c) for c in cams]
        raw = b"".join(bytes(s) for s in structs)
        self.cams = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)
        V, H, W = len(cams), height, width
        self.shape = (V, H, W)
        self.depth = torch.empty((V, H, W), device=device, dtype=torch.float32)
        self.prim_id = torch.empty((V, H, W), device=device, dtype=torch.int32)
        self.obj_id = torch.empty((V, H, W), device=device, dtype=torch.int32)
        self.inst_id = torch.empty((V, H, W), device=device, dtype=torch.int32)
        if srgb:
            self.color = torch.empty((V, H, W), device=device, dtype=torch.int32)
        else:
            self.color = torch.empty((V, H, W, 4), device=device, dtype=torch.float32)
        self.counter = torch.zeros(1, device=device, dtype=torch.int32)
        b = N.vrx_batch()
        b.num_views, b.width, b.height = V, W, H
        b.d_cams = self.cams.data_ptr()
        b.d_depth = self.depth.data_ptr()
        b.d_prim_id = self.prim_id.data_ptr()
        b.d_obj_id = self.obj_id.data_ptr()
        b.d_inst_id = self.inst_id.data_ptr()
        if srgb:
            b.d_color_u32 = self.color.data_ptr()
        else:
            b.d_color_f32 = self.color.data_ptr()
        b.d_work_counter = self.counter.data_ptr()
        b.tile_w, b.tile_h, b.block_threads, b.persistent_blocks = 8, 8, 256, 0
        L = N.hip()
        nbytes = L.vrx_batch_scratch_bytes(C.byref(self.scene), C.byref(self.renderer),
                                           C.byref(b))
        self.scratch = torch.empty(max(int(nbytes), 256), device=device, dtype=torch.uint8)
        b.d_scratch = self.scratch.data_ptr()
        b.scratch_bytes = self.scratch.numel()
        self.b = b
        self.graph = None

    @staticmethod
    def _stream(stream=None) -> int:
        s = stream if stream is not None else torch.cuda.current_stream()
        return s.cuda_stream

    def render(self, stream=None):
        N.check(N.hip().vrx_render_batch(C.byref(self.scene), C.byref(self.renderer),
                                         C.byref(self.b), self._stream(stream)),
                "vrx_render_batch")

    def autotune(self, stream=None) -> float:
        """Times tile shape x workgroup size x LDS stack depth x BVH node width
        (4 only when the scene has 4-wide nodes) on this scene and keeps the
        fastest; returns its time per launch in ms."""
        ms = C.c_float()
        N.check(N.hip().vrx_batch_autotune(C.byref(self.scene), C.byref(self.renderer),
                                           C.byref(self.b), self._stream(stream), C.byref(ms)),
                "vrx_batch_autotune")
        self.release_graph()
        return ms.value

    def capture(self, stream: torch.cuda.Stream):
        """Capture one batched frame (queue reset + persistent launch) into a
        HIP graph on ``stream`` (a non-default stream)."""
        self.release_graph()
        g = C.c_void_p()
        N.check(N.hip().vrx_batch_graph_create(C.byref(self.scene), C.byref(self.renderer),
                                               C.byref(self.b), stream.cuda_stream, C.byref(g)),
                "vrx_batch_graph_create")
        self.graph = g

    def replay(self, stream: torch.cuda.Stream):
        N.check(N.hip().vrx_graph_launch(self.graph, stream.cuda_stream), "vrx_graph_launch")

    def release_graph(self):
        if self.graph:
            N.hip().vrx_graph_destroy(self.graph)
            self.graph = None

    def channels(self) -> dict:
        out = {"depth": self.depth.cpu().numpy(),
               "primitiveId": self.prim_id.cpu().numpy().view(np.uint32),
               "objectId": self.obj_id.cpu().numpy().view(np.uint32),
               "instanceId": self.inst_id.cpu().numpy().view(np.uint32)}
        c = self.color.cpu().numpy()
        out["color"] = c.view(np.uint32) if c.dtype == np.int32 else c
        return out

    def __del__(self):
        try:
            self.release_graph()
        except Exception:
            pass
