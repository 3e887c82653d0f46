"""Raycast / AO frames through the ANARI device on the GPU vs the CPU oracle.

Bar: primitiveId / objectId / instanceId bit-exact; depth and FLOAT32_VEC4
colour within 1e-4 relative (they are expected to be bit-identical: the
device and the oracle use the same fp32 operation order); sRGB8 colour
within 1 code value (GPU powf vs glibc powf).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu


def render_gpu(sc, cam, w, h, kind="raycast", color=A.FLOAT32_VEC4, **rp):
    dev = A.Device()
    s = S.build_anari(dev, sc)
    c = S.make_camera(dev, s, cam)
    r = S.make_renderer(dev, s, kind, **rp)
    f = S.make_frame(dev, s, s.world, c, r, w, h, color=color)
    print(f"[render] {kind} {w}x{h} geoms={len(sc.geoms)} inst={len(sc.instances)} "
          f"vols={len(sc.volumes)}", flush=True)
    dev.render(f)
    dev.ready(f)
    out = {}
    for ch in ["color", "depth", "primitiveId", "objectId", "instanceId"]:
        out[ch], _ = dev.map_frame(f, "channel." + ch)
    out["duration"] = dev.get_property(f, "duration", A.FLOAT32)
    out["rays"] = dev.get_property(f, "raysTraced", A.UINT64)
    out["errors"] = dev.errors()
    s.release()
    dev.close()
    return out


def compare(g, o, exact_color=True):
    assert not g["errors"], g["errors"]
    np.testing.assert_array_equal(g["primitiveId"], o["primitiveId"])
    np.testing.assert_array_equal(g["objectId"], o["objectId"])
    np.testing.assert_array_equal(g["instanceId"], o["instanceId"])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-4, atol=0)
    if g["color"].dtype == np.float32:
        np.testing.assert_allclose(g["color"], o["color"], rtol=1e-4, atol=1e-4)
        frac = np.mean(g["color"] == o["color"])
        print(f"color bit-identical fraction: {frac:.6f}; depth identical "
              f"{np.mean(g['depth'] == o['depth']):.6f}")
    else:
        a = g["color"].view(np.uint8).astype(np.int32)
        b = o["color"].view(np.uint8).astype(np.int32)
        assert np.abs(a - b).max() <= 1


def test_heightfield_raycast_full_size(require_gpu):
    """1024x768 raycast of the 131,072-triangle heightfield."""
    sc, cam = S.heightfield_workload()
    g = render_gpu(sc, cam, 1024, 768)
    o = orc.OracleScene(sc).render(cam, 1024, 768)
    compare(g, o)
    assert (g["depth"] < 1e29).mean() > 0.05


def test_heightfield_srgb(require_gpu):
    sc, cam = S.heightfield_workload()
    g = render_gpu(sc, cam, 256, 192, color=A.UFIXED8_RGBA_SRGB)
    o = orc.OracleScene(sc).render(cam, 256, 192, color_format=orc.ORC_COLOR_SRGB8)
    compare(g, o)


def test_spheres_raycast(require_gpu):
    sc, cam = S.spheres_workload(n=200_000, width=480, height=270, seed=5)
    g = render_gpu(sc, cam, 480, 270)
    o = orc.OracleScene(sc).render(cam, 480, 270)
    compare(g, o)


@pytest.mark.slow
def test_10m_spheres_raycast_full_size(require_gpu):
    sc, cam = S.spheres_workload()
    g = render_gpu(sc, cam, 1920, 1080)
    o = orc.OracleScene(sc).render(cam, 1920, 1080)
    compare(g, o)


def _cylinder_scene(caps):
    rng = np.random.default_rng(4)
    n = 3000
    p0 = rng.random((n, 3), dtype=np.float32)
    p1 = p0 + rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    pos = np.stack([p0, p1], 1).reshape(-1, 3)
    rad = rng.uniform(0.002, 0.01, n).astype(np.float32)
    g = S.GeomDesc("cylinder", pos, radius=rad, caps=caps, surface_id=3,
                   primitive_id=np.arange(n, dtype=np.uint32)[::-1].copy())
    sc = S.SceneDesc(geoms=[g], world_geoms=[0], world_id=9)
    cam = S.CameraDesc((0.5, 0.5, -1.0), (0, 0, 1), (0, 1, 0), math.radians(50), 4 / 3)
    return sc, cam


@pytest.mark.parametrize("caps", [False, True])
def test_cylinders_raycast(require_gpu, caps):
    sc, cam = _cylinder_scene(caps)
    g = render_gpu(sc, cam, 320, 240)
    o = orc.OracleScene(sc).render(cam, 320, 240)
    compare(g, o)


def _instanced_scene():
    """Three meshes, seven instances with rotations (two-level traversal)."""
    geoms = []
    for i, t in enumerate([200, 800, 3000]):
        v, idx = S.gen_sphere_mesh(t, seed=20 + i)
        geoms.append(S.GeomDesc("triangle", v, idx, surface_id=100 + i,
                                color=(0.2 + 0.3 * i, 0.5, 0.7)))
    rng = np.random.default_rng(8)
    insts = []
    for k in range(7):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        s = rng.uniform(0.3, 0.8)
        t = rng.uniform(0, 4, 3)
        M = np.zeros(12, np.float32)
        M[0:9] = (R * s).T.reshape(-1)   # column-major 3x3
        M[9:12] = t
        insts.append(S.InstanceDesc(geoms=[k % 3], xfm=M, id=1000 + k))
    sc = S.SceneDesc(geoms=geoms, instances=insts, world_id=5)
    cam = S.CameraDesc((2.0, 2.0, -6.0), S.look_at((2.0, 2.0, -6.0), (2, 2, 2)), (0, 1, 0),
                       math.radians(50), 1.0)
    return sc, cam


def test_instanced_triangles_raycast(require_gpu):
    sc, cam = _instanced_scene()
    g = render_gpu(sc, cam, 256, 256)
    o = orc.OracleScene(sc).render(cam, 256, 256)
    compare(g, o)
    assert len(np.unique(g["instanceId"])) > 3


def test_mixed_types_transparency_and_culling(require_gpu):
    rng = np.random.default_rng(12)
    c = rng.random((400, 3), dtype=np.float32)
    sph = S.GeomDesc("sphere", c, radius=np.full(400, 0.03, np.float32), surface_id=1,
                     opacity=0.5)
    v, idx = S.gen_heightfield(33, seed=3)
    tri = S.GeomDesc("triangle", v, idx, surface_id=2, cull=True)
    cyl_p = rng.random((200, 3), dtype=np.float32)
    cyl = S.GeomDesc("cylinder", cyl_p, global_radius=0.01, surface_id=4)
    sc = S.SceneDesc(geoms=[sph, tri, cyl], world_geoms=[0, 1, 2], world_id=7)
    cam = S.CameraDesc((0.5, -1.2, 1.0), S.look_at((0.5, -1.2, 1.0), (0.5, 0.5, 0.2)), (0, 0, 1),
                       math.radians(60), 1.5)
    g = render_gpu(sc, cam, 300, 200, background=(0.1, 0.2, 0.3, 1.0))
    o = orc.OracleScene(sc).render(cam, 300, 200, background=(0.1, 0.2, 0.3, 1.0))
    compare(g, o)


def test_ao_renderer_small(require_gpu):
    sc, cam = S.spheres_workload(n=50_000, width=160, height=90, seed=6)
    rp = dict(pixelSamples=4, ambientSamples=2, ambientOcclusionDistance=0.05)
    g = render_gpu(sc, cam, 160, 90, kind="ao", **rp)
    o = orc.OracleScene(sc).render(cam, 160, 90, kind="ao", **rp)
    compare(g, o)


def test_empty_world_background(require_gpu):
    sc = S.SceneDesc(world_id=3)
    cam = S.CameraDesc((0, 0, 0), (0, 0, 1), (0, 1, 0), 1.0, 1.0)
    g = render_gpu(sc, cam, 8, 8, background=(0.25, 0.5, 0.75, 1.0))
    assert (g["primitiveId"] == 0).all()
    assert (g["objectId"] == 0xFFFFFFFF).all() and (g["instanceId"] == 0xFFFFFFFF).all()
    assert (g["depth"] == np.float32(1e30)).all()
    o = orc.OracleScene(sc).render(cam, 8, 8, background=(0.25, 0.5, 0.75, 1.0))
    compare(g, o)


def test_volume_raycast_vs_fp64_oracle(require_gpu):
    """64^3 noise field, EA march at 0.5-voxel steps, early termination at 0.99:
    colour within 1e-4 of the FP64 oracle march; ids and entry depth."""
    data = S.gen_noise_volume(64, seed=4)
    sc, cam = S.volume_workload(n=64, width=200, height=120, data=data)
    g = render_gpu(sc, cam, 200, 120, volumeSamplingRate=1.0)
    o = orc.OracleScene(sc).render(cam, 200, 120, volumeSamplingRate=1.0)
    assert not g["errors"], g["errors"]
    ok = o["ambiguous"] == 0
    assert ok.mean() > 0.99, f"ambiguous fraction {1 - ok.mean():.4f}"
    np.testing.assert_allclose(g["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    np.testing.assert_array_equal(g["primitiveId"], o["primitiveId"])
    np.testing.assert_array_equal(g["objectId"], o["objectId"])
    np.testing.assert_array_equal(g["instanceId"], o["instanceId"])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-5)
    hit = g["depth"] < 1e29
    assert hit.mean() > 0.2 and (g["color"][hit][:, :3].max() > 0.05)


def test_volume_with_surface_in_front_and_inside(require_gpu):
    """A sphere inside the volume: march up to the surface hit, then shade."""
    data = S.gen_noise_volume(48, seed=6)
    sc, cam = S.volume_workload(n=48, width=160, height=100, data=data)
    sc.geoms = [S.GeomDesc("sphere", np.array([[0.5, 0.5, 0.5], [0.2, 0.8, 0.3]], np.float32),
                           radius=np.array([0.25, 0.1], np.float32), surface_id=11)]
    sc.world_geoms = [0]
    g = render_gpu(sc, cam, 160, 100, volumeSamplingRate=0.5)
    o = orc.OracleScene(sc).render(cam, 160, 100, volumeSamplingRate=0.5)
    assert not g["errors"], g["errors"]
    ok = o["ambiguous"] == 0
    assert ok.mean() > 0.99, f"ambiguous fraction {1 - ok.mean():.4f}"
    np.testing.assert_allclose(g["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    np.testing.assert_array_equal(g["primitiveId"], o["primitiveId"])
    np.testing.assert_array_equal(g["objectId"], o["objectId"])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-5)


def _quad_grid(n=33, seed=4):
    """A heightfield as an indexed quad mesh: quad (a, b, c, d) per grid cell."""
    verts, _ = S.gen_heightfield(n, seed)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (i * n + j).reshape(-1)
    quads = np.stack([a, a + 1, a + n + 1, a + n], 1).astype(np.uint32)
    return verts, quads


def test_quads_trace_as_the_reference_triangle_split(require_gpu):
    """Quad geometry (Quad.cpp:124-145): each quad is traced as the triangles
    (x, y, w) and (z, w, y); primitiveId is the triangle index.  The quad
    frame equals the frame of that triangle mesh bit for bit, and the oracle's
    render of it; the soup form (4 vertices per quad) too."""
    verts, quads = _quad_grid()
    cam = S.CameraDesc((0.5, -1.5, 1.2), S.look_at((0.5, -1.5, 1.2), (0.5, 0.5, 0.0)), (0, 0, 1),
                       math.radians(45.0), 4 / 3)
    tris = S.quads_to_triangles(quads, len(verts))
    sc_q = S.SceneDesc(geoms=[S.GeomDesc("quad", verts, quads, surface_id=3)], world_geoms=[0])
    sc_t = S.SceneDesc(geoms=[S.GeomDesc("triangle", verts, tris, surface_id=3)], world_geoms=[0])
    gq = render_gpu(sc_q, cam, 160, 120)
    gt = render_gpu(sc_t, cam, 160, 120)
    o = orc.OracleScene(sc_t).render(cam, 160, 120)
    compare(gq, o)
    for ch in ["color", "depth", "primitiveId", "objectId", "instanceId"]:
        np.testing.assert_array_equal(gq[ch], gt[ch], err_msg=ch)
    hit = gq["depth"] < 1e29
    assert hit.mean() > 0.05 and len(np.unique(gq["primitiveId"][hit])) > 200
    # soup: quad q is vertices 4q .. 4q+3
    vs = verts[quads.reshape(-1)]
    sc_qs = S.SceneDesc(geoms=[S.GeomDesc("quad", vs, surface_id=3)], world_geoms=[0])
    sc_ts = S.SceneDesc(geoms=[S.GeomDesc("triangle", vs, S.quads_to_triangles(None, len(vs)),
                                          surface_id=3)], world_geoms=[0])
    gs = render_gpu(sc_qs, cam, 160, 120)
    compare(gs, orc.OracleScene(sc_ts).render(cam, 160, 120))


def test_quad_soup_must_be_a_multiple_of_four(require_gpu):
    dev = A.Device()
    s = S.AnariScene(dev)
    g = s.keep(dev.new("Geometry", "quad"))
    pos = np.zeros((6, 3), np.float32)
    dev.set(g, "vertex.position", s.keep(dev.array(pos, A.FLOAT32_VEC3)), A.ARRAY1D)
    dev.commit(g)
    w = s.keep(dev.new("World"))
    sf = s.keep(dev.new("Surface"))
    dev.set(sf, "geometry", g, A.GEOMETRY)
    dev.commit(sf)
    dev.set(w, "surface", s.keep(dev.array([sf], A.SURFACE)), A.ARRAY1D)
    dev.commit(w)
    dev.get_property(w, "bounds", A.FLOAT32_BOX3)
    errs = " ".join(str(e) for e in dev.errors())
    assert "non-multiple of 4" in errs, errs
    s.release()
    dev.close()


def _cone_scene(indexed):
    """3,000 capped cones with per-vertex radii (Cone.cpp, Intersectors_ptx.cu:192-246)."""
    rng = np.random.default_rng(6)
    n = 3000
    p0 = rng.random((n, 3), dtype=np.float32)
    p1 = p0 + rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    pos = np.stack([p0, p1], 1).reshape(-1, 3)
    rad = rng.uniform(0.0, 0.015, 2 * n).astype(np.float32)
    idx = None
    if indexed:   # reversed pairs: cone i joins vertices 2i+1 -> 2i
        idx = np.stack([2 * np.arange(n) + 1, 2 * np.arange(n)], 1).astype(np.uint32)
    g = S.GeomDesc("cone", pos, idx, radius=rad, surface_id=5)
    sc = S.SceneDesc(geoms=[g], world_geoms=[0], world_id=2)
    cam = S.CameraDesc((0.5, 0.5, -1.0), (0, 0, 1), (0, 1, 0), math.radians(50), 4 / 3)
    return sc, cam


@pytest.mark.parametrize("indexed", [False, True])
def test_cones_raycast(require_gpu, indexed):
    """Cones vs the oracle: ids bit-exact, depth and colour within 1e-4."""
    sc, cam = _cone_scene(indexed)
    g = render_gpu(sc, cam, 320, 240)
    o = orc.OracleScene(sc).render(cam, 320, 240)
    compare(g, o)
    assert (g["depth"] < 1e29).mean() > 0.05


def test_cones_with_other_types_and_ao(require_gpu):
    """A world mixing cones, quads and spheres (generic kernel) under the AO
    renderer, against the oracle."""
    sc_c, cam = _cone_scene(False)
    verts, quads = _quad_grid(17, seed=8)
    c, r = S.gen_spheres(2000, seed=3, rmin=2e-3, rmax=6e-3)
    geoms = [sc_c.geoms[0], S.GeomDesc("quad", verts * 0.5 + 0.25, quads, surface_id=7),
             S.GeomDesc("sphere", c, radius=r, surface_id=8)]
    sc = S.SceneDesc(geoms=geoms, world_geoms=[0, 1, 2], world_id=1)
    tris = S.quads_to_triangles(quads, len(verts))
    sc_o = S.SceneDesc(geoms=[geoms[0], S.GeomDesc("triangle", verts * 0.5 + 0.25, tris, surface_id=7),
                              geoms[2]], world_geoms=[0, 1, 2], world_id=1)
    rp = dict(pixelSamples=2, ambientSamples=2, ambientOcclusionDistance=0.05)
    g = render_gpu(sc, cam, 160, 120, kind="ao", **rp)
    o = orc.OracleScene(sc_o).render(cam, 160, 120, kind="ao", **rp)
    compare(g, o)


def _curve_scene(indexed):
    """400 random polylines of 8 segments (round linear curves, Curve.cpp)
    with per-vertex radii; or, as a soup, segments 0 .. nv/2 - 1."""
    rng = np.random.default_rng(9)
    lines, per = 400, 9
    start = rng.random((lines, 1, 3), dtype=np.float32)
    steps = rng.normal(0, 0.02, (lines, per, 3)).astype(np.float32)
    pos = (start + np.cumsum(steps, 1)).reshape(-1, 3)
    rad = rng.uniform(0.001, 0.008, len(pos)).astype(np.float32)
    idx = None
    if indexed:   # each polyline's 8 segments start at its first 8 vertices
        idx = (np.arange(lines)[:, None] * per + np.arange(per - 1)[None]).reshape(-1).astype(np.uint32)
    g = S.GeomDesc("curve", pos, idx, radius=rad, surface_id=4)
    sc = S.SceneDesc(geoms=[g], world_geoms=[0], world_id=6)
    cam = S.CameraDesc((0.5, 0.5, -1.0), (0, 0, 1), (0, 1, 0), math.radians(50), 4 / 3)
    return sc, cam


def _segment_t(g, s, org, d):
    a = int(g.index[s]) if g.index is not None else s
    t = np.zeros(1, np.float32)
    n = np.zeros(3, np.float32)
    keep = [np.ascontiguousarray(v, np.float32) for v in (org, d, g.position[a], g.position[a + 1])]
    hit = orc.lib().orc_intersect_round_cone(*[k.ctypes.data for k in keep], float(g.radius[a]),
                                             float(g.radius[a + 1]), t.ctypes.data, n.ctypes.data)
    return float(t[0]) if hit else None


@pytest.mark.parametrize("indexed", [False, True])
def test_curves_raycast(require_gpu, indexed):
    """Round linear curves vs the oracle: depth/colour within 1e-4, ids
    bit-exact except where two adjacent segments meet the ray within 4 ulps
    of each other (the shared joint sphere seen through two formulas); there
    either id is a correct closest hit."""
    sc, cam = _curve_scene(indexed)
    g = render_gpu(sc, cam, 320, 240)
    o = orc.OracleScene(sc).render(cam, 320, 240)
    diff = np.argwhere(g["primitiveId"] != o["primitiveId"])
    geom = sc.geoms[0]
    for (y, x) in diff:
        pg, po = int(g["primitiveId"][y, x]), int(o["primitiveId"][y, x])
        org, d = orc.camera_ray(cam, 320, 240, float(x), float(y))
        tg, to = _segment_t(geom, pg, org, d), _segment_t(geom, po, org, d)
        assert abs(pg - po) == 1 and tg is not None and to is not None, (x, y, pg, po)
        assert abs(tg - to) <= 4 * np.spacing(np.float32(to)), (x, y, tg, to)
        g["primitiveId"][y, x] = po
    assert len(diff) <= 4
    compare(g, o)
    assert (g["depth"] < 1e29).mean() > 0.02


def test_curve_index_out_of_range_is_rejected(require_gpu):
    dev = A.Device()
    s = S.AnariScene(dev)
    g = s.keep(dev.new("Geometry", "curve"))
    dev.set(g, "vertex.position", s.keep(dev.array(np.zeros((4, 3), np.float32), A.FLOAT32_VEC3)),
            A.ARRAY1D)
    dev.set(g, "primitive.index", s.keep(dev.array(np.array([0, 3], np.uint32), A.UINT32)),
            A.ARRAY1D)
    dev.commit(g)
    sf = s.keep(dev.new("Surface"))
    dev.set(sf, "geometry", g, A.GEOMETRY)
    dev.commit(sf)
    w = s.keep(dev.new("World"))
    dev.set(w, "surface", s.keep(dev.array([sf], A.SURFACE)), A.ARRAY1D)
    dev.commit(w)
    dev.get_property(w, "bounds", A.FLOAT32_BOX3)   # flushes the deferred commits
    errs = " ".join(str(e) for e in dev.errors())
    assert "needs vertex 4" in errs, errs
    s.release()
    dev.close()


def test_transform_array_million_instances(require_gpu):
    """A million instances from one transform array, raycast and AO, against
    the oracle (which walks the entries through its own instance BVH)."""
    sc, cam = S.instanced_spheres_workload(width=320, height=240)
    g = render_gpu(sc, cam, 320, 240)
    os_ = orc.OracleScene(sc)
    o = os_.render(cam, 320, 240)
    compare(g, o)
    hit = g["depth"] < 1e29
    assert hit.mean() > 0.5
    assert ((g["instanceId"][hit] - 7) % 3 == 0).all()
    rp = dict(pixelSamples=1, ambientSamples=2, ambientOcclusionDistance=0.02)
    g = render_gpu(sc, cam, 160, 120, kind="ao", **rp)
    o = os_.render(cam, 160, 120, kind="ao", **rp)
    compare(g, o)
