"""ANARI object semantics on the GPU that a drop-in must keep:

* object arrays: a Group or World holds its own references to the objects of
  its surface / instance arrays, and an array changed through map/unmap is
  seen without re-committing its users (helium ChangeObserverPtr; reference
  Group.h:90-109, World.h:73-77, Instance.h:60-64, array/Array.cpp:134-141);
* Array1D regions (`begin` / `end`, array/Array1D.cpp:43-91);
* arrays over device memory are used in place (array/Array.cpp:38-65);
* a frame completion callback may call the API while another thread waits
  for the frame (frame/Frame.cu:295-304, 430-438).

Frames are compared with the CPU oracle at the bar of test_gpu_render.py:
ids exact, depth and FLOAT32 colour within 1e-4.
"""
import ctypes as C
import math
import os
import subprocess
import threading

import numpy as np
import pytest

from oracle import oracle as orc
from tests.devmem import DeviceBuffer
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)
W, H = 96, 72
CAM = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), math.radians(45.0), W / H)


def _frame_out(dev, f):
    out = {}
    for ch in ["color", "depth", "primitiveId", "objectId", "instanceId"]:
        out[ch], _ = dev.map_frame(f, "channel." + ch)
    return out


def _compare(g, o):
    np.testing.assert_array_equal(g["primitiveId"], o["primitiveId"])
    np.testing.assert_array_equal(g["objectId"], o["objectId"])
    np.testing.assert_array_equal(g["instanceId"], o["instanceId"])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-4, atol=0)
    np.testing.assert_allclose(g["color"], o["color"], rtol=1e-4, atol=1e-4)


def _spheres(seed, n=4000, sid=1):
    c, r = S.gen_spheres(n, seed=seed, rmin=3e-3, rmax=1.2e-2)
    return S.GeomDesc("sphere", c, radius=r, surface_id=sid)


def _own_surface(dev, g):
    """A surface whose handles (geometry, arrays, material) the caller owns."""
    s = S.AnariScene(dev)
    surf = S._surface(dev, s, g)
    return surf, s


def test_surface_array_remapped_after_its_handles_are_released(require_gpu):
    """World and group keep their surfaces alive: the app maps the surface
    array, writes a new handle, unmaps, releases every handle of the old
    surface, then commits an unrelated object and renders; the frame is the
    oracle's frame of the new surface."""
    dev = A.Device()
    ga, gb = _spheres(11, sid=1), _spheres(12, sid=2)
    surf_a, own_a = _own_surface(dev, ga)
    surf_b, own_b = _own_surface(dev, gb)
    keep = S.AnariScene(dev)
    arr = keep.keep(dev.array([surf_a], A.SURFACE))
    world = keep.keep(dev.new("World"))
    dev.set(world, "surface", arr, A.ARRAY1D)
    dev.commit(world)
    cam = S.make_camera(dev, keep, CAM)
    ren = S.make_renderer(dev, keep, "raycast")
    f = S.make_frame(dev, keep, world, cam, ren, W, H)
    print("[api] frame 1: surface A", flush=True)
    dev.render(f)
    dev.ready(f)
    _compare(_frame_out(dev, f), orc.OracleScene(S.SceneDesc(geoms=[ga], world_geoms=[0]))
             .render(CAM, W, H))
    ptr = dev.L.anariMapArray(dev.h, arr)
    C.memmove(ptr, C.byref(C.c_void_p(surf_b)), 8)
    dev.L.anariUnmapArray(dev.h, arr)
    own_a.release()                       # the old surface is now held only internally
    light = keep.keep(dev.new("Light", "directional"))
    dev.commit(light)                     # any commit: the flush must not touch freed objects
    print("[api] frame 2: surface B after remap", flush=True)
    dev.render(f)
    dev.ready(f)
    _compare(_frame_out(dev, f), orc.OracleScene(S.SceneDesc(geoms=[gb], world_geoms=[0]))
             .render(CAM, W, H))
    assert not dev.errors(), dev.errors()
    own_b.release()
    keep.release()
    dev.close()


def _instance_transform_scene(dev, xf, ids, keep):
    g = S.GeomDesc("sphere", np.zeros((1, 3), np.float32), radius=np.ones(1, np.float32),
                   surface_id=5)
    surf = S._surface(dev, keep, g)
    grp = keep.keep(dev.new("Group"))
    dev.set(grp, "surface", keep.keep(dev.array([surf], A.SURFACE)), A.ARRAY1D)
    dev.commit(grp)
    m4 = np.zeros((len(xf), 4, 4), np.float32)
    m4[:, :, :3] = xf.reshape(-1, 4, 3)
    m4[:, 3, 3] = 1.0
    xarr = keep.keep(dev.array(m4, A.FLOAT32_MAT4))
    inst = keep.keep(dev.new("Instance", "transform"))
    dev.set(inst, "group", grp, A.GROUP)
    dev.set(inst, "transform", xarr, A.ARRAY1D)
    dev.set(inst, "id", keep.keep(dev.array(ids, A.UINT32)), A.ARRAY1D)
    dev.commit(inst)
    world = keep.keep(dev.new("World"))
    dev.set(world, "instance", keep.keep(dev.array([inst], A.INSTANCE)), A.ARRAY1D)
    dev.commit(world)
    return world, xarr, g


def _xforms(seed, n):
    rng = np.random.default_rng(seed)
    xf = np.zeros((n, 12), np.float32)
    sc = rng.uniform(0.03, 0.08, n).astype(np.float32)
    xf[:, 0] = xf[:, 4] = xf[:, 8] = sc
    xf[:, 9:12] = rng.random((n, 3), dtype=np.float32)
    return xf


def test_transform_array_edit_without_recommit(require_gpu):
    """Instance transform array (Instance.cpp:115-128) edited through
    map/unmap; the instance is not committed again and the next frame uses
    the new transforms."""
    n = 40
    xf0, xf1 = _xforms(3, n), _xforms(4, n)
    ids = np.arange(n, dtype=np.uint32) + 100
    dev = A.Device()
    keep = S.AnariScene(dev)
    world, xarr, g = _instance_transform_scene(dev, xf0, ids, keep)
    cam = S.make_camera(dev, keep, CAM)
    ren = S.make_renderer(dev, keep, "raycast")
    f = S.make_frame(dev, keep, world, cam, ren, W, H)
    for k, xf in enumerate([xf0, xf1]):
        if k:
            m4 = np.zeros((n, 4, 4), np.float32)
            m4[:, :, :3] = xf.reshape(-1, 4, 3)
            m4[:, 3, 3] = 1.0
            ptr = dev.L.anariMapArray(dev.h, xarr)
            C.memmove(ptr, m4.ctypes.data, m4.nbytes)
            dev.L.anariUnmapArray(dev.h, xarr)
        print(f"[api] transform array version {k}", flush=True)
        dev.render(f)
        dev.ready(f)
        sc = S.SceneDesc(geoms=[g], instances=[S.InstanceDesc(geoms=[0], xfm=xf, ids=ids)])
        _compare(_frame_out(dev, f), orc.OracleScene(sc).render(CAM, W, H))
    assert not dev.errors(), dev.errors()
    keep.release()
    dev.close()


def test_transfer_function_edit_without_recommit(require_gpu):
    """The volume's opacity array edited through map/unmap: the LUT is
    rebuilt (TransferFunction1D.cpp:101-150) without re-committing the
    volume, and the frame matches the FP64 oracle march of the new LUT."""
    data = S.gen_noise_volume(32, seed=8)
    sc, cam = S.volume_workload(n=32, width=80, height=60, data=data)
    v = sc.volumes[0]
    dev = A.Device()
    keep = S.AnariScene(dev)
    field_ = keep.keep(dev.new("SpatialField", "structuredRegular"))
    dev.set(field_, "data", keep.keep(dev.array(v.data, A.FLOAT32, dims=(32, 32, 32))), A.ARRAY3D)
    dev.set(field_, "origin", v.origin, A.FLOAT32_VEC3)
    dev.set(field_, "spacing", v.spacing, A.FLOAT32_VEC3)
    dev.commit(field_)
    vol = keep.keep(dev.new("Volume", "transferFunction1D"))
    dev.set(vol, "value", field_, A.SPATIAL_FIELD)
    dev.set(vol, "color", keep.keep(dev.array(v.color, A.FLOAT32_VEC3)), A.ARRAY1D)
    op_arr = keep.keep(dev.array(v.opacity, A.FLOAT32))
    dev.set(vol, "opacity", op_arr, A.ARRAY1D)
    dev.set(vol, "valueRange", v.value_range, A.FLOAT32_VEC2)
    dev.set(vol, "unitDistance", v.unit_distance, A.FLOAT32)
    dev.set(vol, "id", v.id, A.UINT32)
    dev.commit(vol)
    world = keep.keep(dev.new("World"))
    dev.set(world, "volume", keep.keep(dev.array([vol], A.VOLUME)), A.ARRAY1D)
    dev.set(world, "id", sc.world_id, A.UINT32)
    dev.commit(world)
    c = S.make_camera(dev, keep, cam)
    r = S.make_renderer(dev, keep, "raycast", volumeSamplingRate=1.0)
    f = S.make_frame(dev, keep, world, c, r, 80, 60)
    for k, op in enumerate([v.opacity, np.array([0.2, 0.02], np.float32)]):
        if k:
            ptr = dev.L.anariMapArray(dev.h, op_arr)
            C.memmove(ptr, op.ctypes.data, op.nbytes)
            dev.L.anariUnmapArray(dev.h, op_arr)
        print(f"[api] opacity version {k}", flush=True)
        dev.render(f)
        dev.ready(f)
        g = _frame_out(dev, f)
        v.opacity = op
        o = orc.OracleScene(sc).render(cam, 80, 60, volumeSamplingRate=1.0)
        ok = o["ambiguous"] == 0
        assert ok.mean() > 0.98
        np.testing.assert_allclose(g["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
        np.testing.assert_array_equal(g["objectId"], o["objectId"])
    assert not dev.errors(), dev.errors()
    keep.release()
    dev.close()


def test_array1d_region(require_gpu):
    """Array1D `begin`/`end` (Array1D.cpp:43-91): the geometry sees only
    elements [begin, end) of its arrays, primitive indices count from begin;
    committing a new region re-commits the geometry (its observer)."""
    g = _spheres(21, n=6000, sid=4)
    dev = A.Device()
    keep = S.AnariScene(dev)
    pos = keep.keep(dev.array(g.position, A.FLOAT32_VEC3))
    rad = keep.keep(dev.array(g.radius, A.FLOAT32))
    geom = keep.keep(dev.new("Geometry", "sphere"))
    dev.set(geom, "vertex.position", pos, A.ARRAY1D)
    dev.set(geom, "vertex.radius", rad, A.ARRAY1D)
    dev.commit(geom)
    surf = keep.keep(dev.new("Surface"))
    dev.set(surf, "geometry", geom, A.GEOMETRY)
    dev.set(surf, "id", 4, A.UINT32)
    dev.commit(surf)
    world = keep.keep(dev.new("World"))
    dev.set(world, "surface", keep.keep(dev.array([surf], A.SURFACE)), A.ARRAY1D)
    dev.commit(world)
    cam = S.make_camera(dev, keep, CAM)
    ren = S.make_renderer(dev, keep, "raycast")
    f = S.make_frame(dev, keep, world, cam, ren, W, H)
    for b, e in [(1000, 4500), (3000, 6000), (0, 6000)]:
        for a in (pos, rad):
            dev.set(a, "begin", b, A.UINT64)
            dev.set(a, "end", e, A.UINT64)
            dev.commit(a)
        print(f"[api] region [{b}, {e})", flush=True)
        dev.render(f)
        dev.ready(f)
        sub = S.GeomDesc("sphere", g.position[b:e], radius=g.radius[b:e], surface_id=4)
        _compare(_frame_out(dev, f),
                 orc.OracleScene(S.SceneDesc(geoms=[sub], world_geoms=[0])).render(CAM, W, H))
    assert not dev.errors(), dev.errors()
    keep.release()
    dev.close()


def test_device_memory_arrays_used_in_place(require_gpu):
    """Shared arrays over device memory (Array.cpp:38-65) render the same
    frame as host arrays; the device reads them in place, so a write to
    the device memory followed by unmap is seen by the next frame.  A
    captured device array calls its deleter on release."""
    g = _spheres(31, n=5000, sid=7)
    g2 = _spheres(32, n=5000, sid=7)
    dpos = DeviceBuffer.from_array(g.position)
    drad = DeviceBuffer.from_array(g.radius)
    freed = []
    DELETER = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)

    @DELETER
    def on_delete(user, mem):
        freed.append((user, mem))

    dev = A.Device()
    keep = S.AnariScene(dev)
    pos = keep.keep(dev.shared_array(dpos.addr, A.FLOAT32_VEC3, len(g.position)))
    rad = dev.L.anariNewArray1D(dev.h, C.c_void_p(drad.addr), C.cast(on_delete, C.c_void_p),
                                C.c_void_p(77), A.FLOAT32, len(g.radius))
    assert rad
    geom = keep.keep(dev.new("Geometry", "sphere"))
    dev.set(geom, "vertex.position", pos, A.ARRAY1D)
    dev.set(geom, "vertex.radius", rad, A.ARRAY1D)
    dev.commit(geom)
    dev.release(rad)          # the geometry keeps the captured array alive
    surf = keep.keep(dev.new("Surface"))
    dev.set(surf, "geometry", geom, A.GEOMETRY)
    dev.set(surf, "id", 7, A.UINT32)
    dev.commit(surf)
    world = keep.keep(dev.new("World"))
    dev.set(world, "surface", keep.keep(dev.array([surf], A.SURFACE)), A.ARRAY1D)
    dev.commit(world)
    cam = S.make_camera(dev, keep, CAM)
    ren = S.make_renderer(dev, keep, "raycast")
    f = S.make_frame(dev, keep, world, cam, ren, W, H)
    print("[api] device-memory arrays, frame 1", flush=True)
    dev.render(f)
    dev.ready(f)
    _compare(_frame_out(dev, f),
             orc.OracleScene(S.SceneDesc(geoms=[g], world_geoms=[0])).render(CAM, W, H))
    ptr = dev.L.anariMapArray(dev.h, pos)
    assert ptr == dpos.addr, "a device array maps to the application's device memory"
    hip = dpos.L
    assert hip.vrx_memcpy_sync(C.c_void_p(dpos.addr), C.c_void_p(g2.position.ctypes.data),
                               C.c_size_t(g2.position.nbytes), 0) == 0
    dev.L.anariUnmapArray(dev.h, pos)
    print("[api] device-memory arrays, frame 2 (positions rewritten on the device)", flush=True)
    dev.render(f)
    dev.ready(f)
    moved = S.GeomDesc("sphere", g2.position, radius=g.radius, surface_id=7)
    _compare(_frame_out(dev, f),
             orc.OracleScene(S.SceneDesc(geoms=[moved], world_geoms=[0])).render(CAM, W, H))
    assert not dev.errors(), dev.errors()
    assert freed == []
    keep.release()            # geometry released -> captured array's deleter runs
    assert len(freed) == 1 and freed[0][0] == 77 and freed[0][1] == drad.addr, freed
    dev.close()
    dpos.free()
    drad.free()


def test_callback_calls_the_api_while_another_thread_waits(require_gpu):
    """The completion callback calls anariGetProperty (a blocking world
    query and a frame property) while the main thread sits in
    anariFrameReady(WAIT) and a second thread creates objects; nothing
    deadlocks, WAIT returns after the callback has run, once per frame."""
    dev = A.Device()
    keep = S.AnariScene(dev)
    sc = S.SceneDesc(geoms=[_spheres(41, sid=3)], world_geoms=[0])
    s = S.build_anari(dev, sc)
    cam = S.make_camera(dev, keep, CAM)
    ren = S.make_renderer(dev, keep, "ao", pixelSamples=2)
    f = S.make_frame(dev, keep, s.world, cam, ren, W, H)
    seen = []

    @CB
    def on_done(user, device, frame):
        box = (C.c_float * 6)()
        ok = dev.L.anariGetProperty(device, s.world, b"bounds", A.FLOAT32_BOX3, box,
                                    C.sizeof(box), A.WAIT)
        ns = C.c_int32()
        ok2 = dev.L.anariGetProperty(device, frame, b"numSamples", A.INT32, C.byref(ns), 4,
                                     A.NO_WAIT)
        seen.append((ok, ok2, list(box)))

    dev.set(f, "frameCompletionCallback", C.cast(on_done, C.c_void_p).value,
            A.FRAME_COMPLETION_CALLBACK)
    dev.commit(f)
    stop = threading.Event()
    made = []

    def creator():
        while not stop.is_set():
            h = dev.L.anariNewCamera(dev.h, b"perspective")
            made.append(h)
            dev.L.anariRelease(dev.h, h)

    t = threading.Thread(target=creator)
    t.start()
    try:
        for k in range(4):
            print(f"[api] frame {k} with an API-calling callback", flush=True)
            dev.render(f)
            assert dev.ready(f, wait=True)
            assert len(seen) == k + 1, seen
            assert dev.ready(f, wait=False)
    finally:
        stop.set()
        t.join()
    assert all(ok and ok2 for ok, ok2, _ in seen), seen
    lo, hi = np.array(seen[0][2][:3]), np.array(seen[0][2][3:])
    assert np.all(lo < hi)
    assert made and not dev.errors(), dev.errors()
    s.release()
    keep.release()
    dev.close()


API_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "api")


@pytest.mark.parametrize("prog", ["device_init", "mt_object_creation"])
def test_reference_api_programs(require_gpu, prog):
    """The reference's tests/api programs (TestDeviceInit.cpp:37-44,
    TestMultiThreadedObjectCreation.cpp:57-95) as C++ clients of this
    device: forceInit on the device handle; 4 threads x 100 x {light,
    camera, triangle geometry, world, frame} created and released."""
    exe = os.path.join(API_DIR, prog)
    print(f"[api] {exe}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "errors=0" in r.stdout
