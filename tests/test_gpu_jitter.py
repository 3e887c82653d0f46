"""Volume jitter on the GPU (renderer parameter volumeJitter, the two
per-segment jitters of volumeIntegration.h:83,120) against the oracle drawing
the same PCG32 numbers: the raycast frame of a volume alone (which then takes
the generic kernel), a volume around a sphere under the AO renderer (jitter
draws interleaved with camera and AO draws), shadow rays through a volume
under directLight, and the batched launch against per-view frames.

Bars as without jitter: ids exact, entry depth within 1e-5 (the depth is the
segment entry, before jitter), colour within 1e-4 on pixels the FP64 oracle
does not flag.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import anari as A
from visrtx_amd import scenes as S
from tests.test_gpu_render import render_gpu

pytestmark = pytest.mark.gpu


def _check(g, o, min_ok=0.98):
    assert not g["errors"], g["errors"]
    ok = o["ambiguous"] == 0
    assert ok.mean() > min_ok, f"ambiguous fraction {1 - ok.mean():.4f}"
    np.testing.assert_allclose(g["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    for ch in ("primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(g[ch], o[ch])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-5)


def test_jittered_volume_raycast(require_gpu):
    data = S.gen_noise_volume(48, seed=14)
    sc, cam = S.volume_workload(n=48, width=160, height=120, data=data)
    rp = dict(volumeSamplingRate=1.0)
    print("[jitter] volume raycast 160x120", flush=True)
    g = render_gpu(sc, cam, 160, 120, volumeJitter=True, **rp)
    o = orc.OracleScene(sc).render(cam, 160, 120, volumeJitter=True, **rp)
    _check(g, o)
    plain = orc.OracleScene(sc).render(cam, 160, 120, **rp)
    hit = g["depth"] < 1e29
    assert hit.mean() > 0.2
    # jitter moves the samples: the colour changes, the entry depth does not
    assert (np.abs(plain["color"] - o["color"]).max(-1)[hit] > 1e-4).mean() > 0.5
    np.testing.assert_array_equal(plain["depth"], o["depth"])
    g2 = render_gpu(sc, cam, 160, 120, volumeJitter=True, **rp)
    np.testing.assert_array_equal(g["color"], g2["color"])


def test_jittered_volume_around_a_sphere_ao(require_gpu):
    data = S.gen_noise_volume(40, seed=15)
    sc, cam = S.volume_workload(n=40, width=128, height=96, data=data)
    sc.geoms = [S.GeomDesc("sphere", np.array([[0.5, 0.5, 0.5]], np.float32),
                           radius=np.array([0.2], np.float32), surface_id=9)]
    sc.world_geoms = [0]
    rp = dict(volumeSamplingRate=0.5, pixelSamples=2, ambientSamples=2,
              ambientOcclusionDistance=0.3)
    print("[jitter] AO 2 spp through a volume, 128x96", flush=True)
    g = render_gpu(sc, cam, 128, 96, kind="ao", volumeJitter=True, **rp)
    o = orc.OracleScene(sc).render(cam, 128, 96, kind="ao", volumeJitter=True, **rp)
    _check(g, o, min_ok=0.97)


def test_jittered_volume_shadows_direct_light(require_gpu):
    data = S.gen_noise_volume(40, seed=2)
    vol = S.VolumeDesc(data, origin=(-1.0, 0.6, -1.0), spacing=(2.0 / 39, 0.5 / 39, 2.0 / 39),
                       color=np.array([[0.2, 0.4, 1.0], [1.0, 0.6, 0.2]], np.float32),
                       opacity=np.array([0.0, 0.6], np.float32), value_range=(0.0, 1.0), id=40)
    plane = np.array([[-2, 0, -2], [2, 0, -2], [2, 0, 2], [-2, 0, 2]], np.float32)
    g0 = S.GeomDesc("triangle", plane, np.array([[0, 2, 1], [0, 3, 2]], np.uint32),
                    color=(0.9, 0.8, 0.7), surface_id=1)
    sc = S.SceneDesc(geoms=[g0], world_geoms=[0], volumes=[vol], world_id=0,
                     lights=[S.LightDesc("directional", direction=(0.2, -1.0, 0.1),
                                         irradiance=2.0)])
    cam = S.CameraDesc((0.0, 3.0, -2.5), S.look_at((0.0, 3.0, -2.5), (0.0, 0.0, 0.0)),
                       (0, 1, 0), np.radians(45.0), 4 / 3)
    rp = dict(pixelSamples=1, ambientSamples=0, volumeSamplingRate=0.5)
    print("[jitter] directLight volume shadows 128x96", flush=True)
    g = render_gpu(sc, cam, 128, 96, kind="directLight", volumeJitter=True, **rp)
    o = orc.OracleScene(sc).render(cam, 128, 96, kind="directLight", volumeJitter=True, **rp)
    _check(g, o, min_ok=0.95)
    plain = orc.OracleScene(sc).render(cam, 128, 96, kind="directLight", **rp)
    assert (np.abs(plain["color"] - o["color"]).max(-1) > 1e-4).mean() > 0.05


def test_jittered_batch_equals_per_view_frames(require_gpu):
    torch = pytest.importorskip("torch")
    from visrtx_amd import batch as B
    data = S.gen_noise_volume(32, seed=16)
    sc, cam = S.volume_workload(n=32, width=64, height=48, data=data)
    cams = []
    for k in range(3):
        pos = (1.9 - 0.4 * k, 1.5, -1.0 + 0.3 * k)
        cams.append(S.CameraDesc(pos, S.look_at(pos, (0.5, 0.5, 0.5)), (0, 1, 0),
                                 np.radians(35.0), 64 / 48))
    dev = A.Device()
    s = S.build_anari(dev, sc)
    r = S.make_renderer(dev, s, "raycast", volumeSamplingRate=1.0, volumeJitter=True)
    bf = B.BatchFrame(B.scene_of(dev, s.world), B.renderer_of(dev, r), cams, 64, 48)
    assert bf.renderer.volume_jitter == 1
    print("[jitter] batched 3 views 64x48", flush=True)
    bf.render()
    torch.cuda.synchronize()
    got = bf.channels()
    for v, c in enumerate(cams):
        cc = S.make_camera(dev, s, c)
        f = S.make_frame(dev, s, s.world, cc, r, 64, 48)
        dev.render(f)
        dev.ready(f)
        for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
            np.testing.assert_array_equal(got[ch][v], dev.map_frame(f, "channel." + ch)[0],
                                          err_msg=f"view {v}: {ch}")
        s.drop(f)
        s.drop(cc)
    assert not dev.errors(), dev.errors()
    bf = None
    s.release()
    dev.close()
