"""The oracle's directLight renderer (DirectLight_ptx.cu:71-420) against
closed forms, and its opacity-accumulating shadow/AO any-hit
(DirectLight_ptx.cu:255-280, AmbientOcclusion_ptx.cu:53-69).

Scenes are a lit plane y = 0 seen from above.  Occluders face downwards and
the renderer culls back-facing triangles, so the camera sees through them
while shadow and AO rays (never culled) hit them.  Light values follow
MatteShader_ptx.cu:71-79: base/pi * max(0, N.L) * radiance, with no distance
falloff; a point-light value at a pixel is checked against the closed form
bracketed over the pixel footprint (the directLight camera rays are jittered
inside the pixel)."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S

W = H = 48
BASE = (0.9, 0.6, 0.3)
INV_PI = 1.0 / math.pi


def _quad(y, half, down, cx=0.0, cz=0.0, opacity=1.0, sid=5, color=(0.5, 0.5, 0.5)):
    """Axis-aligned square at height y as two triangles; normal -y if down."""
    v = np.array([[cx - half, y, cz - half], [cx + half, y, cz - half],
                  [cx + half, y, cz + half], [cx - half, y, cz + half]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]] if not down else [[0, 1, 2], [0, 2, 3]], np.uint32)
    return S.GeomDesc("triangle", v, idx, surface_id=sid, color=color, opacity=opacity)


def _plane_scene(occluders=(), lights=(), instances=()):
    geoms = [_quad(0.0, 3.0, False, sid=1, color=BASE)] + list(occluders)
    sc = S.SceneDesc(geoms=geoms, world_geoms=list(range(len(geoms))), world_id=0,
                     lights=list(lights), instances=list(instances))
    cam = S.CameraDesc((0.0, 5.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0),
                       math.radians(40.0), 1.0)
    return sc, cam


def _render(sc, cam, **rp):
    rp.setdefault("ambientSamples", 0)
    rp.setdefault("cullTriangleBackfaces", 1)
    return orc.OracleScene(sc).render(cam, W, H, kind="directLight", **rp)


def _plane_points(cam, ox=0.5, oy=0.5):
    """Plane hit points of the rays through (px + ox, py + oy)."""
    pts = np.zeros((H, W, 3))
    for py in range(H):
        for px in range(W):
            o, d = orc.camera_ray(cam, W, H, px + ox, py + oy)
            t = -float(o[1]) / float(d[1])
            pts[py, px] = np.asarray(o, np.float64) + t * np.asarray(d, np.float64)
    return pts


def _bracket(cam, f, foot):
    """Bounds of f over each pixel's footprint for f decreasing with the
    distance from foot (x, z): the camera looks straight down, so a footprint
    is an axis-aligned rectangle; f is largest at the rectangle point nearest
    the foot and smallest at the farthest corner."""
    corners = np.stack([_plane_points(cam, ox, oy) for ox in (0.0, 1.0) for oy in (0.0, 1.0)])
    lo_xz = corners[..., [0, 2]].min(0)
    hi_xz = corners[..., [0, 2]].max(0)
    near = np.clip(np.asarray(foot, np.float64), lo_xz, hi_xz)
    far = np.where(np.abs(lo_xz - foot) > np.abs(hi_xz - foot), lo_xz, hi_xz)

    def at(xz):
        return f(np.stack([xz[..., 0], np.zeros(xz.shape[:-1]), xz[..., 1]], -1))

    return at(far), at(near)


def test_directional_light_constant_on_plane():
    L = S.LightDesc("directional", color=(1.0, 0.5, 0.25), direction=(0.3, -1.0, 0.2),
                    irradiance=2.0)
    sc, cam = _plane_scene(lights=[L])
    o = _render(sc, cam)
    d = np.array(L.direction) / np.linalg.norm(L.direction)
    expect = np.array(BASE) * INV_PI * (-d[1]) * 2.0 * np.array(L.color)
    np.testing.assert_allclose(o["color"][..., :3], np.broadcast_to(expect, (H, W, 3)),
                               rtol=2e-6)
    assert (o["objectId"] == 1).all()


def test_ambient_term_and_default_radiance():
    """ambientColor * ambientRadiance * tint; directLight's ambientRadiance
    defaults to 0 (DirectLight.cpp:46), so an unlit scene is black."""
    sc, cam = _plane_scene()
    o = _render(sc, cam, ambientRadiance=0.5, ambientColor=(1.0, 0.8, 0.6))
    expect = np.array([1.0, 0.8, 0.6]) * 0.5 * np.array(BASE)
    np.testing.assert_allclose(o["color"][..., :3], np.broadcast_to(expect, (H, W, 3)),
                               rtol=2e-6)
    o0 = _render(sc, cam)
    assert (o0["color"][..., :3] == 0).all() and (o0["color"][..., 3] == 1).all()
    od = orc.OracleScene(sc).render(cam, W, H, kind="default", ambientSamples=0)
    np.testing.assert_array_equal(od["color"], o0["color"])


def test_point_light_no_falloff():
    """radiance = color * intensity at any distance (sampleLight.h:70-79)."""
    h = 6.0
    L = S.LightDesc("point", position=(0.4, h, -0.3), intensity=3.0)
    sc, cam = _plane_scene(lights=[L])
    o = _render(sc, cam)

    def f(p):
        v = np.array(L.position) - p
        return 3.0 * INV_PI * v[..., 1] / np.linalg.norm(v, axis=-1)

    lo, hi = _bracket(cam, f, (0.4, -0.3))
    assert (hi - lo).max() < 1e-2 * hi.max()
    for ch in range(3):
        c = o["color"][..., ch] / BASE[ch]
        assert (c >= lo * (1 - 1e-5)).all() and (c <= hi * (1 + 1e-5)).all()


def test_spot_light_smoothstep():
    """Inside the inner cone full intensity, outside the opening angle zero,
    smoothstep between (sampleLight.h:81-100)."""
    L = S.LightDesc("spot", position=(0.0, 0.8, 0.0), direction=(0.0, -1.0, 0.0),
                    intensity=2.0, opening_angle=math.radians(60.0), falloff_angle=0.2)
    sc, cam = _plane_scene(lights=[L])
    o = _render(sc, cam)
    co, ci = math.cos(math.radians(60.0)), math.cos(math.radians(60.0) - 0.4)

    def f(p):
        v = np.array(L.position) - p
        c = v[..., 1] / np.linalg.norm(v, axis=-1)   # = N.L = cos to the axis
        s = np.clip((c - co) / (ci - co), 0.0, 1.0)
        return 2.0 * INV_PI * c * (s * s * (3.0 - 2.0 * s))

    lo, hi = _bracket(cam, f, (0.0, 0.0))
    c = o["color"][..., 0] / BASE[0]
    assert (c >= lo - 1e-6).all() and (c <= hi + 1e-6).all()
    assert (c == 0).mean() > 0.2 and (c > 0.5).mean() > 0.05


def test_shadow_ray_length():
    """A point light's shadow ray ends at the light; a spot light's ends one
    unit from the surface (the reference takes the length of the normalised
    direction, sampleLight.h:81-82), so an occluder 2 units up shadows the
    point light but not the spot light at the same place."""
    occ = _quad(2.0, 0.2, True)
    pl = S.LightDesc("point", position=(0.0, 3.0, 0.0), intensity=1.0)
    sp = S.LightDesc("spot", position=(0.0, 3.0, 0.0), direction=(0.0, -1.0, 0.0),
                     intensity=1.0, opening_angle=math.radians(170.0), falloff_angle=0.01)
    pts = _plane_points(cam=_plane_scene()[1])
    under = (np.abs(pts[..., 0]) < 0.3) & (np.abs(pts[..., 2]) < 0.3)
    assert under.sum() > 20
    sc, cam = _plane_scene(occluders=[occ], lights=[pl])
    o = _render(sc, cam)
    assert (o["color"][under][:, :3] == 0).all()
    far = (np.abs(pts[..., 0]) > 1.0) | (np.abs(pts[..., 2]) > 1.0)
    assert (o["color"][far][:, 0] > 0).all()
    beyond = _quad(4.0, 0.2, True)            # above the point light: no shadow
    sc, cam = _plane_scene(occluders=[beyond], lights=[pl])
    assert (_render(sc, cam)["color"][under][:, 0] > 0).all()
    sc, cam = _plane_scene(occluders=[occ], lights=[sp])
    assert (_render(sc, cam)["color"][under][:, 0] > 0).all()


@pytest.mark.parametrize("ops,expect", [((0.5,), 0.5), ((0.5, 0.6), 0.2),
                                        ((0.9, 0.85), 0.015), ((0.9, 0.95), 0.0),
                                        ((0.995,), 0.0), ((1.0,), 0.0)])
def test_translucent_shadows(ops, expect):
    """Shadow transmission 1 - o, o accumulated over every occluder hit,
    o += a (1 - o); once o reaches 0.99 the ray counts as blocked."""
    L = S.LightDesc("directional", direction=(0.0, -1.0, 0.0), irradiance=1.0)
    occ = [_quad(1.0 + 0.5 * k, 10.0, True, opacity=a, sid=10 + k) for k, a in enumerate(ops)]
    sc, cam = _plane_scene(occluders=occ, lights=[L])
    o = _render(sc, cam)
    lit = np.array(BASE) * INV_PI
    np.testing.assert_allclose(o["color"][..., :3], np.broadcast_to(lit * expect, (H, W, 3)),
                               rtol=1e-5, atol=1e-7)


def test_surface_attenuation_counts_reports():
    """A sphere reports one intersection (its near root), a cylinder two (in
    and out, Intersectors_ptx.cu:163-169), each adding its opacity."""
    sph = S.GeomDesc("sphere", np.array([[0, 0, 5]], np.float32),
                     radius=np.array([1.0], np.float32), opacity=0.5)
    cyl = S.GeomDesc("cylinder", np.array([[-3, 0, 5], [3, 0, 5]], np.float32),
                     radius=np.array([1.0], np.float32), opacity=0.5)
    for g, expect in [(sph, 0.5), (cyl, 0.75)]:
        s = orc.OracleScene(S.SceneDesc(geoms=[g], world_geoms=[0]))
        a = s.surface_attenuation((0, 0, 0), (0, 0, 1))
        assert abs(a - expect) < 1e-6, (g.type, a)
        assert s.surface_attenuation((0, 0, 0), (0, 0, 1), 0.0, 3.0) == 0.0
    opaque = orc.OracleScene(S.SceneDesc(geoms=[_quad(1.0, 1.0, True)], world_geoms=[0]))
    assert opaque.surface_attenuation((0, 0, 0), (0, 1, 0)) == 1.0


def test_ao_counts_translucent_occluders():
    """AO rays accumulate opacity like shadow rays: under a canopy of opacity
    a the AO factor is 1 - a (AmbientOcclusion_ptx.cu:53-69)."""
    a = 0.4
    sc, cam = _plane_scene(occluders=[_quad(0.5, 1e4, True, opacity=a)])
    for kind in ("ao", "directLight"):
        o = orc.OracleScene(sc).render(cam, W, H, kind=kind, ambientSamples=4, ambientRadiance=1.0,
                                       cullTriangleBackfaces=1)
        np.testing.assert_allclose(o["color"][..., :3],
                                   np.broadcast_to((1 - a) * np.array(BASE), (H, W, 3)),
                                   rtol=2e-4)


def test_translucent_surface_accumulates_onto_itself():
    """DirectLight_ptx.cu:372-381 replaces the layer's colour/opacity by the
    shading result and accumulates it again (colour c (2 - a), opacity
    a (2 - a)), then multiplies by that opacity: a surface of opacity a and
    shaded colour c over an opaque black background shows c a (2 - a)^2."""
    a = 0.5
    L = S.LightDesc("directional", direction=(0.0, -1.0, 0.0), irradiance=1.0)
    sc, cam = _plane_scene(lights=[L])
    sc.geoms[0] = _quad(0.0, 3.0, False, sid=1, color=BASE, opacity=a)
    o = _render(sc, cam)
    c = np.array(BASE) * INV_PI
    np.testing.assert_allclose(o["color"][..., :3],
                               np.broadcast_to(c * a * (2 - a) ** 2, (H, W, 3)), rtol=1e-5)
    np.testing.assert_allclose(o["color"][..., 3], 1.0)


def test_volume_shadow_closed_form():
    """A homogeneous slab between a directional light and the plane: the
    shadow ray's transmission is (1 - alpha)^(n * step / unitDistance) over its
    n samples (volumeIntegration.h:64-103 without colour)."""
    # slab y in [1, 2.04]; march step = min(spacing) / 2 = 0.1, so neither the
    # primary chords (about 10.4 steps) nor the shadow chord sit on a sample
    nx, ny, nz = 61, 5, 9
    data = np.full((nz, ny, nx), 0.5, np.float32)
    alpha = 0.3
    vol = S.VolumeDesc(data, origin=(-6.0, 1.0, -6.0), spacing=(0.2, 0.26, 1.5),
                       color=np.ones((2, 3), np.float32),
                       opacity=np.full(2, alpha, np.float32), value_range=(0.0, 1.0))
    d = np.array([0.3, -1.0, 0.0])
    L = S.LightDesc("directional", direction=tuple(d), irradiance=1.0)
    sc, cam = _plane_scene(lights=[L])
    sc.volumes = [vol]
    o = orc.OracleScene(sc).render(cam, W, H, kind="directLight", ambientSamples=0,
                                   volumeSamplingRate=1.0, cullTriangleBackfaces=1)
    ok = o["ambiguous"] == 0                    # primary rays march the slab too
    assert ok.mean() > 0.95
    u = -d / np.linalg.norm(d)
    chord = np.float32(4 * np.float32(0.26)) / u[1]
    step = 0.2 / 2.0
    n = math.floor(chord / step + 1e-9) + 1
    assert abs(chord / step - round(chord / step)) > 0.05
    trans = (1.0 - alpha) ** (n * step)
    lit = np.array(BASE) * INV_PI * u[1]
    np.testing.assert_allclose(o["color"][ok][:, :3],
                               np.broadcast_to(lit * trans, (int(ok.sum()), 3)), rtol=1e-5)


def test_instance_light_transform():
    """A group light under an instance transform: a point light at the origin
    of a group translated by (0, 5, 0) equals a world point light at (0, 5, 0);
    a directional light is rotated by the instance's linear part."""
    pl = S.LightDesc("point", position=(0.0, 0.0, 0.0), intensity=2.0)
    xfm = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0.0, 5.0, 0.0], np.float32)
    sc_i, cam = _plane_scene(instances=[S.InstanceDesc(geoms=[], xfm=xfm, lights=[pl])])
    sc_w, _ = _plane_scene(lights=[S.LightDesc("point", position=(0.0, 5.0, 0.0), intensity=2.0)])
    np.testing.assert_array_equal(_render(sc_i, cam)["color"], _render(sc_w, cam)["color"])
    # rotation by 90 degrees about z maps direction (1, 0, 0) to (0, 1, 0):
    # the light travels along -y after the instance transform of (-1, 0, 0)
    rz = np.array([0, 1, 0, -1, 0, 0, 0, 0, 1, 0, 0, 0], np.float32)
    dl = S.LightDesc("directional", direction=(-1.0, 0.0, 0.0), irradiance=1.0)
    sc_r, _ = _plane_scene(instances=[S.InstanceDesc(geoms=[], xfm=rz, lights=[dl])])
    o = _render(sc_r, cam)
    np.testing.assert_allclose(o["color"][..., :3],
                               np.broadcast_to(np.array(BASE) * INV_PI, (H, W, 3)), rtol=1e-5)


def test_directional_light_under_scaling_instance_is_not_renormalised():
    """xfmVec of the instance's matrix is applied to the direction with no
    renormalisation (sampleLight.h:58-66): under a uniform scale of 2 the
    Lambert term doubles."""
    dl = S.LightDesc("directional", direction=(0.0, -1.0, 0.0), irradiance=1.0)
    x2 = np.array([2, 0, 0, 0, 2, 0, 0, 0, 2, 0, 0, 0], np.float32)
    sc, cam = _plane_scene(instances=[S.InstanceDesc(geoms=[], xfm=x2, lights=[dl])])
    o = _render(sc, cam)
    np.testing.assert_allclose(o["color"][..., :3],
                               np.broadcast_to(2.0 * np.array(BASE) * INV_PI, (H, W, 3)),
                               rtol=1e-5)


def test_spot_light_under_instance_moves_but_keeps_its_axis():
    """A spot light's position follows the instance transform, its axis does
    not (sampleLight.h:81-86 reads ld.spot.direction untransformed): a group
    spot at the origin pointing down, under a translation (0, 0.8, 0) and a
    90-degree rotation about z, lights the plane as a world spot at
    (0, 0.8, 0) pointing down."""
    sp = S.LightDesc("spot", position=(0.0, 0.0, 0.0), direction=(0.0, -1.0, 0.0),
                     intensity=2.0, opening_angle=math.radians(60.0), falloff_angle=0.2)
    rz = np.array([0, 1, 0, -1, 0, 0, 0, 0, 1, 0.0, 0.8, 0.0], np.float32)
    sc_i, cam = _plane_scene(instances=[S.InstanceDesc(geoms=[], xfm=rz, lights=[sp])])
    ws = S.LightDesc("spot", position=(0.0, 0.8, 0.0), direction=(0.0, -1.0, 0.0),
                     intensity=2.0, opening_angle=math.radians(60.0), falloff_angle=0.2)
    sc_w, _ = _plane_scene(lights=[ws])
    np.testing.assert_array_equal(_render(sc_i, cam)["color"], _render(sc_w, cam)["color"])
