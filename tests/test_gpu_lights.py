"""directLight renderer on the GPU vs the CPU oracle (DirectLight_ptx.cu,
gpu/sampleLight.h, MatteShader_ptx.cu), and the opacity-accumulating
shadow / AO any-hit in both renderers that use it.

Bar as for the other renderers: ids bit-exact, depth and FLOAT32_VEC4 colour
within 1e-4 (the device and the oracle share the fp32 operation order; the
any-hit sums translucent opacities in traversal order, which differs between
the two by rounding only).  Volume shadows compare outside the pixels the
FP64 oracle flags as fp32-ambiguous."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_gpu_render import compare, render_gpu
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu


def _quad(y, half, down, opacity=1.0, sid=5, color=(0.5, 0.5, 0.5), cx=0.0, cz=0.0):
    v = np.array([[cx - half, y, cz - half], [cx + half, y, cz - half],
                  [cx + half, y, cz + half], [cx - half, y, cz + half]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]] if down else [[0, 2, 1], [0, 3, 2]], np.uint32)
    return S.GeomDesc("triangle", v, idx, surface_id=sid, color=color, opacity=opacity)


def _lit_scene(seed=3):
    """A plane, 3000 spheres, 200 cylinders, a translucent canopy, two
    instances of a sphere mesh whose groups carry a point light and a
    directional + spot pair, and world directional, point and spot lights."""
    rng = np.random.default_rng(seed)
    n = 3000
    c = np.stack([rng.uniform(-1, 1, n), rng.uniform(0.05, 1.0, n),
                  rng.uniform(-1, 1, n)], 1).astype(np.float32)
    r = rng.uniform(0.01, 0.04, n).astype(np.float32)
    m = 200
    p0 = np.stack([rng.uniform(-1, 1, m), rng.uniform(0.0, 0.6, m),
                   rng.uniform(-1, 1, m)], 1).astype(np.float32)
    p1 = p0 + rng.normal(0, 0.15, (m, 3)).astype(np.float32)
    cyl = np.stack([p0, p1], 1).reshape(-1, 3)
    mv, mi = S.gen_sphere_mesh(600, seed=seed)
    geoms = [_quad(0.0, 3.0, False, sid=1, color=(0.8, 0.8, 0.75)),
             S.GeomDesc("sphere", c, radius=r, surface_id=2, color=(0.9, 0.4, 0.2)),
             S.GeomDesc("cylinder", cyl, radius=rng.uniform(0.005, 0.02, m).astype(np.float32),
                        surface_id=3, color=(0.2, 0.5, 0.9)),
             _quad(1.3, 0.8, True, opacity=0.5, sid=4, color=(0.3, 0.9, 0.3), cx=0.3),
             S.GeomDesc("triangle", mv, mi, surface_id=6, color=(0.7, 0.7, 0.2))]
    pl = S.LightDesc("point", color=(1.0, 0.9, 0.7), position=(0.0, 2.5, 0.0), intensity=0.6)
    t = math.radians(30.0)
    rot = np.array([math.cos(t), 0, -math.sin(t), 0, 1, 0, math.sin(t), 0, math.cos(t)],
                   np.float32) * 0.3
    insts = [S.InstanceDesc(geoms=[4], xfm=np.concatenate([rot, [0.5, 0.3, -0.4]]).astype(
                 np.float32), id=100, lights=[pl]),
             S.InstanceDesc(geoms=[4], xfm=np.concatenate([rot, [-0.6, 0.35, 0.5]]).astype(
                 np.float32), id=101, lights=[
                     # under the instance's 0.3 scale: a directional light's
                     # direction is not renormalised, a spot's axis is not
                     # transformed (sampleLight.h:58-100)
                     S.LightDesc("directional", direction=(0.2, -1.0, 0.1), irradiance=1.0),
                     S.LightDesc("spot", position=(0.0, 3.0, 0.0), direction=(0.1, -1.0, 0.0),
                                 intensity=0.9, opening_angle=math.radians(70.0),
                                 falloff_angle=0.2)])]
    lights = [S.LightDesc("directional", color=(1.0, 0.95, 0.9), direction=(0.4, -1.0, 0.25),
                          irradiance=1.5),
              S.LightDesc("point", position=(-0.8, 1.2, -0.6), intensity=0.8),
              S.LightDesc("spot", position=(0.6, 0.9, 0.2), direction=(-0.3, -1.0, 0.0),
                          intensity=1.2, opening_angle=math.radians(50.0), falloff_angle=0.15)]
    sc = S.SceneDesc(geoms=geoms, world_geoms=[0, 1, 2, 3], instances=insts, world_id=7,
                     lights=lights)
    cam = S.CameraDesc((0.0, 2.2, -2.6), S.look_at((0.0, 2.2, -2.6), (0.0, 0.3, 0.0)),
                       (0, 1, 0), math.radians(45.0), 4 / 3)
    return sc, cam


RP = dict(pixelSamples=2, ambientSamples=2, ambientRadiance=0.3,
          ambientOcclusionDistance=0.3)


def test_direct_light_mixed_scene(require_gpu):
    sc, cam = _lit_scene()
    g = render_gpu(sc, cam, 160, 120, kind="directLight", **RP)
    o = orc.OracleScene(sc).render(cam, 160, 120, kind="directLight", **RP)
    compare(g, o)
    # every light type and the instance light contribute
    assert len(np.unique(g["objectId"])) >= 6
    assert g["color"][..., :3].max() > 0.3


def test_default_renderer_is_direct_light(require_gpu):
    sc, cam = _lit_scene(seed=5)
    rp = dict(pixelSamples=1, ambientSamples=1)
    a = render_gpu(sc, cam, 96, 72, kind="default", **rp)
    b = render_gpu(sc, cam, 96, 72, kind="directLight", **rp)
    for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(a[ch], b[ch], err_msg=ch)


def test_ao_renderer_translucent_occluders(require_gpu):
    """AO rays accumulate the opacity of translucent surfaces
    (AmbientOcclusion_ptx.cu:53-69), on the GPU as in the oracle."""
    sc, cam = _lit_scene(seed=7)
    sc.geoms[3] = _quad(0.4, 2.0, True, opacity=0.35, sid=4)
    rp = dict(pixelSamples=2, ambientSamples=4, ambientOcclusionDistance=1.0)
    g = render_gpu(sc, cam, 128, 96, kind="ao", **rp)
    o = orc.OracleScene(sc).render(cam, 128, 96, kind="ao", **rp)
    compare(g, o)


def test_direct_light_volume_shadows(require_gpu):
    """Shadow rays through a noise volume (volumeAttenuation,
    DirectLight_ptx.cu:62-69) onto a plane and spheres."""
    data = S.gen_noise_volume(40, seed=2)
    vol = S.VolumeDesc(data, origin=(-1.0, 0.6, -1.0), spacing=(2.0 / 39, 0.5 / 39, 2.0 / 39),
                       color=np.array([[0.2, 0.4, 1.0], [1.0, 0.6, 0.2]], np.float32),
                       opacity=np.array([0.0, 0.6], np.float32), value_range=(0.0, 1.0), id=40)
    sc, cam = _lit_scene(seed=11)
    sc.geoms, sc.world_geoms, sc.instances = sc.geoms[:3], [0, 1, 2], []
    sc.volumes = [vol]
    rp = dict(pixelSamples=1, ambientSamples=0, volumeSamplingRate=0.5)
    g = render_gpu(sc, cam, 128, 96, kind="directLight", **rp)
    o = orc.OracleScene(sc).render(cam, 128, 96, kind="directLight", **rp)
    assert not g["errors"], g["errors"]
    ok = o["ambiguous"] == 0
    assert ok.mean() > 0.97, f"ambiguous fraction {1 - ok.mean():.4f}"
    np.testing.assert_allclose(g["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    np.testing.assert_array_equal(g["primitiveId"], o["primitiveId"])
    np.testing.assert_array_equal(g["objectId"], o["objectId"])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-5)
    # the volume darkens part of the plane: lit pixels differ from a render
    # without it
    sc.volumes = []
    o2 = orc.OracleScene(sc).render(cam, 128, 96, kind="directLight", **rp)
    assert (np.abs(o2["color"] - o["color"]).max(-1) > 1e-3).mean() > 0.05


def test_unsupported_light_is_skipped(require_gpu):
    """An hdri light (not on this device) is reported and skipped; the other
    lights render as without it."""
    sc, cam = _lit_scene(seed=13)
    sc.instances = []
    dev = A.Device()
    s = S.build_anari(dev, sc)
    hd = s.keep(dev.new("Light", "hdri"))
    dev.commit(hd)
    la = s.keep(dev.array([S._light(dev, s, L) for L in sc.lights] + [hd], A.LIGHT))
    dev.set(s.world, "light", la, A.ARRAY1D)
    dev.commit(s.world)
    c = S.make_camera(dev, s, cam)
    r = S.make_renderer(dev, s, "directLight", pixelSamples=1, ambientSamples=0)
    f = S.make_frame(dev, s, s.world, c, r, 64, 48)
    dev.render(f)
    dev.ready(f)
    col, _ = dev.map_frame(f, "channel.color")
    col = col.copy()
    s.release()
    dev.close()
    o = orc.OracleScene(sc).render(cam, 64, 48, kind="directLight", pixelSamples=1,
                                   ambientSamples=0)
    np.testing.assert_allclose(col, o["color"], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("prim", ["sphere", "triangle", "cylinder"])
def test_direct_light_primitive_specialised_kernels(require_gpu, prim):
    """Worlds of one primitive type take the kernel specialised for it (no
    volume code); directLight there against the oracle."""
    lights = [S.LightDesc("directional", direction=(0.3, -1.0, 0.4), irradiance=1.5),
              S.LightDesc("point", position=(0.3, 1.4, -0.2), intensity=0.7),
              S.LightDesc("spot", position=(0.5, 0.9, 0.5), direction=(0.0, -1.0, 0.0),
                          intensity=1.0, opening_angle=math.radians(80.0), falloff_angle=0.3)]
    if prim == "sphere":
        sc, cam = S.spheres_workload(n=100_000, width=160, height=120, seed=8)
    elif prim == "triangle":
        sc, cam = S.heightfield_workload()
        cam = S.CameraDesc(cam.position, cam.direction, cam.up, cam.fovy, 160 / 120)
    else:
        rng = np.random.default_rng(5)
        p0 = rng.random((4000, 3), dtype=np.float32)
        p1 = p0 + rng.normal(0, 0.05, (4000, 3)).astype(np.float32)
        sc = S.SceneDesc(geoms=[S.GeomDesc("cylinder", np.stack([p0, p1], 1).reshape(-1, 3),
                                           radius=rng.uniform(0.003, 0.01, 4000).astype(
                                               np.float32), caps=True, surface_id=2)],
                         world_geoms=[0], world_id=3)
        cam = S.CameraDesc((0.5, 0.5, -1.0), (0, 0, 1), (0, 1, 0), math.radians(50), 4 / 3)
    sc.lights = lights
    rp = dict(pixelSamples=2, ambientSamples=1, ambientRadiance=0.2,
              ambientOcclusionDistance=0.05)
    g = render_gpu(sc, cam, 160, 120, kind="directLight", **rp)
    o = orc.OracleScene(sc).render(cam, 160, 120, kind="directLight", **rp)
    compare(g, o)
    assert (g["color"][..., :3] > 0).any()
