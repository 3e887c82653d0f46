"""Frame semantics through the ANARI API on the GPU: the completion callback
fires once per rendered frame with that frame's handle (Frame.cu:295-304),
frameReady(NO_WAIT) turns true once the frame is done, accumulation resets on
a camera commit and the raycast renderer stops at one sample (Raycast.cpp:
45-50), ray counters and the traversal stack-overflow check stay clean."""
import ctypes as C
import math

import numpy as np
import pytest

from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)


def _scene(dev):
    c, r = S.gen_spheres(20_000, seed=5, rmin=2e-3, rmax=8e-3)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", c, radius=r, surface_id=1)], world_geoms=[0])
    s = S.build_anari(dev, sc)
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), math.radians(45.0), 1.0)
    return s, cam


def test_completion_callbacks_and_ready(require_gpu):
    dev = A.Device()
    s, cam = _scene(dev)
    r = S.make_renderer(dev, s, "ao", pixelSamples=1, ambientSamples=1)
    seen = []

    @CB
    def on_done(user, device, frame):
        seen.append((user, frame))

    frames = []
    for k in range(3):
        c = S.make_camera(dev, s, cam)
        f = S.make_frame(dev, s, s.world, c, r, 64, 64)
        dev.set(f, "frameCompletionCallback", C.cast(on_done, C.c_void_p).value,
                A.FRAME_COMPLETION_CALLBACK)
        dev.set(f, "frameCompletionCallbackUserData", 1000 + k, A.VOID_POINTER)
        dev.commit(f)
        frames.append(f)
    print("[frame] 3 frames with callbacks", flush=True)
    for f in frames:
        dev.render(f)
    for f in frames:
        assert dev.ready(f, wait=True)
        assert dev.ready(f, wait=False)
    got = sorted((u, fr) for u, fr in seen)
    assert got == sorted((1000 + k, frames[k]) for k in range(3)), got
    for f in frames:
        assert dev.get_property(f, "raysTraced", A.UINT64) >= 64 * 64
    assert not dev.errors(), dev.errors()
    s.release()
    dev.close()


def test_accumulation_and_sample_limit(require_gpu):
    dev = A.Device()
    s, cam = _scene(dev)
    c = S.make_camera(dev, s, cam)
    ao = S.make_renderer(dev, s, "ao", pixelSamples=1, ambientSamples=2, sampleLimit=3)
    f = S.make_frame(dev, s, s.world, c, ao, 48, 48)
    ns = []
    for _ in range(5):
        dev.render(f)
        dev.ready(f)
        ns.append(dev.get_property(f, "numSamples", A.INT32))
    # frameID counts accumulated samples; once it reaches sampleLimit the
    # frame is not rendered again (Frame.cu:251-253)
    assert ns == [0, 1, 2, 3, 3], ns
    dev.commit(c)                       # a camera commit resets accumulation
    dev.render(f)
    dev.ready(f)
    assert dev.get_property(f, "numSamples", A.INT32) == 0
    rc = S.make_renderer(dev, s, "raycast")
    dev.set(f, "renderer", rc, A.RENDERER)
    dev.commit(f)
    a = None
    for _ in range(3):
        dev.render(f)
        dev.ready(f)
        b = dev.map_frame(f, "channel.color")[0].copy()
        if a is not None:
            np.testing.assert_array_equal(a, b)   # single-shot renderer
        a = b
    assert not dev.errors(), dev.errors()
    s.release()
    dev.close()


def test_launch_autotuning_keeps_frames_identical(require_gpu):
    """The frame times node width x LDS stack depth on its first frames (on a
    one-level scene of one primitive type, then the fastest of those with the
    long-ray pass: five candidates) and keeps the fastest; every frame's
    channels stay bit-identical."""
    dev = A.Device()
    s, cam = _scene(dev)
    c = S.make_camera(dev, s, cam)
    r = S.make_renderer(dev, s, "raycast")
    f = S.make_frame(dev, s, s.world, c, r, 96, 80)
    outs = []
    for _ in range(10):
        dev.commit(c)
        dev.render(f)
        dev.ready(f)
        outs.append({ch: dev.map_frame(f, "channel." + ch)[0].copy()
                     for ch in ("color", "depth", "primitiveId")})
    for o in outs[1:]:
        for ch in o:
            np.testing.assert_array_equal(o[ch], outs[0][ch], err_msg=ch)
    assert dev.get_property(f, "tunedNodeWidth", A.INT32) in (2, 4)
    assert dev.get_property(f, "tunedStackDepth", A.INT32) in (16, 32)
    assert dev.get_property(f, "tunedLongRaySteps", A.INT32) in (0, 32)
    # an explicit parameter is honoured, not tuned
    dev.set(r, "nodeWidth", 2, A.INT32)
    dev.commit(r)
    for _ in range(6):
        dev.commit(c)                   # raycast frames render once per camera commit
        dev.render(f)
        dev.ready(f)
    assert dev.get_property(f, "tunedNodeWidth", A.INT32) == 2
    assert not dev.errors(), dev.errors()
    s.release()
    dev.close()


def test_group_bounds_property(require_gpu):
    """Group "bounds" (Group.cpp:65-86): union of the surfaces' BVH boxes and
    the volumes' boxes in object space, available before any world uses the
    group."""
    dev = A.Device()
    c, r = S.gen_spheres(5000, seed=9, rmin=2e-3, rmax=9e-3)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", c, radius=r, surface_id=1)], world_geoms=[0])
    s = S.build_anari(dev, sc)
    surf = S._surface(dev, s, sc.geoms[0])
    vol = S._volume(dev, s, S.VolumeDesc(np.zeros((4, 5, 6), np.float32),
                                         origin=(-1.0, 0.25, 0.5), spacing=(0.5, 0.5, 0.5)))
    grp = s.keep(dev.new("Group"))
    dev.set(grp, "surface", s.keep(dev.array([surf], A.SURFACE)), A.ARRAY1D)
    dev.commit(grp)
    got = np.array(dev.get_property(grp, "bounds", A.FLOAT32_BOX3), np.float32)
    want = np.concatenate([(c - r[:, None]).min(0), (c + r[:, None]).max(0)]).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    dev.set(grp, "volume", s.keep(dev.array([vol], A.VOLUME)), A.ARRAY1D)
    dev.commit(grp)
    got = np.array(dev.get_property(grp, "bounds", A.FLOAT32_BOX3), np.float32)
    vlo, vhi = np.array([-1.0, 0.25, 0.5]), np.array([-1.0, 0.25, 0.5]) + 0.5 * np.array([5, 4, 3])
    np.testing.assert_allclose(got, np.concatenate([np.minimum(want[:3], vlo),
                                                    np.maximum(want[3:], vhi)]), rtol=1e-6)
    assert not dev.errors(), dev.errors()
    s.release()
    dev.close()
