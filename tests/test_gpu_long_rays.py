"""The raycast renderer's long-ray passes (renderer parameter longRaySteps):
primary walks still running after N steps are queued and finished by a
second launch in which 16 lanes walk 16 t-segments of each queued ray and
keep the closest hit by take()'s rule.  Every intersection test is the
same, so frames are bit-identical to the single walk for any N, on the
three primitive-specialised kernels, with translucent layers (the layer loop
restarts in the second pass) and under the frame autotuner; the rays-traced
counter is unchanged.  The heightfield and translucent cases are also
checked against the oracle (Raycast_ptx.cu:60-179)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_gpu_render import compare, render_gpu
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

CHANNELS = ("color", "depth", "primitiveId", "objectId", "instanceId")


def _same(a, b):
    assert not a["errors"] and not b["errors"], (a["errors"], b["errors"])
    for ch in CHANNELS:
        np.testing.assert_array_equal(a[ch], b[ch], err_msg=ch)
    assert a["rays"] == b["rays"]


@pytest.mark.parametrize("width", [2, 4])
def test_heightfield_long_rays_bit_identical(require_gpu, width):
    """The full 1024x768 heightfield frame (the rays along the horizon are
    the ones deferred), against the frame without the pass and the oracle."""
    sc, cam = S.heightfield_workload()
    base = render_gpu(sc, cam, 1024, 768, longRaySteps=0, nodeWidth=width)
    for n in (1, 6, 32):
        print(f"[long] heightfield width {width} steps {n}", flush=True)
        _same(render_gpu(sc, cam, 1024, 768, longRaySteps=n, nodeWidth=width), base)
    o = orc.OracleScene(sc).render(cam, 1024, 768)
    compare(base, o)


def test_spheres_and_cylinders_long_rays(require_gpu):
    c, r = S.gen_spheres(60000, seed=31, rmin=1e-3, rmax=6e-3)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", c, radius=r, surface_id=2)], world_geoms=[0],
                     world_id=0)
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), 0.8, 4 / 3)
    base = render_gpu(sc, cam, 320, 240, longRaySteps=0)
    print("[long] spheres", flush=True)
    _same(render_gpu(sc, cam, 320, 240, longRaySteps=3), base)
    rng = np.random.default_rng(32)
    p0 = rng.random((3000, 3)).astype(np.float32)
    p1 = (p0 + 0.05 * rng.standard_normal((3000, 3))).astype(np.float32)
    pos = np.stack([p0, p1], axis=1).reshape(-1, 3)
    sc = S.SceneDesc(geoms=[S.GeomDesc("cylinder", pos, global_radius=0.004, caps=True,
                                       surface_id=4)], world_geoms=[0], world_id=0)
    base = render_gpu(sc, cam, 320, 240, longRaySteps=0)
    print("[long] cylinders", flush=True)
    _same(render_gpu(sc, cam, 320, 240, longRaySteps=2), base)


def test_translucent_layers_long_rays(require_gpu):
    """Surface opacity 0.3: each pixel walks several layers; a deferral in a
    later layer re-renders the whole pixel in the second pass."""
    c, r = S.gen_spheres(20000, seed=33, rmin=2e-3, rmax=1e-2)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", c, radius=r, surface_id=2, opacity=0.3)],
                     world_geoms=[0], world_id=0)
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), 0.8, 4 / 3)
    base = render_gpu(sc, cam, 256, 192, longRaySteps=0)
    for n in (2, 12):
        print(f"[long] translucent steps {n}", flush=True)
        _same(render_gpu(sc, cam, 256, 192, longRaySteps=n), base)
    o = orc.OracleScene(sc).render(cam, 256, 192)
    compare(base, o)


def test_autotuner_frames(require_gpu):
    """With longRaySteps unset the frame autotuner times each candidate
    launch (node width x LDS stack, then the fastest with the long-ray
    pass) on its first frames: every frame is the same frame, equal to the
    one without the pass, and the tuner reports its choice."""
    sc, cam = S.heightfield_workload(width=256, height=192)
    ref = render_gpu(sc, cam, 256, 192, longRaySteps=0)
    dev = A.Device()
    s = S.build_anari(dev, sc)
    c = S.make_camera(dev, s, cam)
    r = S.make_renderer(dev, s, "raycast")
    f = S.make_frame(dev, s, s.world, c, r, 256, 192, color=A.FLOAT32_VEC4)
    print("[long] autotuner frames", flush=True)
    for i in range(10):
        dev.commit(c)   # a camera commit starts a new frame (the raycast renderer is single-shot)
        dev.render(f)
        dev.ready(f)
        for ch in CHANNELS:
            np.testing.assert_array_equal(dev.map_frame(f, "channel." + ch)[0], ref[ch],
                                          err_msg=f"frame {i} {ch}")
    assert dev.get_property(f, "tunedLongRaySteps", A.INT32) in (0, 32)
    assert not dev.errors(), dev.errors()
    s.release()
    dev.close()


@pytest.mark.parametrize("stack", [8, 64])
def test_edge_cases_long_rays(require_gpu, stack):
    """One primitive (a root leaf), two primitives, and a world with no
    surfaces, at a budget of one step and at both ends of the LDS stack
    depth: the same frames as without the pass."""
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), 0.8, 4 / 3)
    for n in (1, 2):
        c = np.array([[0.5, 0.5, 0.5], [0.55, 0.5, 0.6]], np.float32)[:n]
        r = np.array([0.2, 0.15], np.float32)[:n]
        sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", c, radius=r, surface_id=2)], world_geoms=[0],
                         world_id=0)
        base = render_gpu(sc, cam, 64, 48, longRaySteps=0, stackDepthLDS=stack)
        print(f"[long] {n} sphere(s), stack {stack}", flush=True)
        _same(render_gpu(sc, cam, 64, 48, longRaySteps=1, stackDepthLDS=stack), base)
        assert (base["depth"] < 1e29).any()
    sc = S.SceneDesc(world_id=0)
    base = render_gpu(sc, cam, 32, 24, longRaySteps=0)
    _same(render_gpu(sc, cam, 32, 24, longRaySteps=1), base)
