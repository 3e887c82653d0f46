"""Pins the CPU oracle's intersectors, traversal, camera, channel packing and
volume march (oracle/oracle.c) against independent references:

- sphere: FP64 closed-form near root; rays starting inside report nothing
  (only the near root is reported, Intersectors_ptx.cu:94-98);
- triangle: FP64 Moller-Trumbore (hit/miss and t), watertightness on a fan;
- cylinder: FP64 quadratic for the infinite cylinder clipped to the caps;
- BVH closest hit == brute force over every primitive, including the
  (t, entry, primitive) tie rule on duplicated primitives;
- camera: geometric invariants (centre ray, corner ray, field of view);
- packing: glm packUnorm4x8 / sRGB known answers;
- volume: constant fields (closed-form transmittance) and linear fields
  (trilinear reproduces them exactly).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_sphere_matches_fp64_near_root():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        c = rng.uniform(-1, 1, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.05, 1.0))
        d = _unit(c - o + rng.normal(0, 0.4, 3)).astype(np.float32)
        t = np.zeros(1, np.float32)
        hit = orc.lib().orc_intersect_sphere(o.ctypes.data, d.ctypes.data, c.ctypes.data,
                                             float(r), t.ctypes.data)
        oc = o.astype(np.float64) - c
        b = np.dot(oc, d.astype(np.float64))
        disc = b * b - (np.dot(oc, oc) - float(r) ** 2)
        if disc > 1e-6:
            assert hit
            t_ref = -b - math.sqrt(disc)
            assert abs(t[0] - t_ref) <= 2e-5 * max(1.0, abs(t_ref))
        elif disc < -1e-6:
            assert not hit


def test_sphere_from_inside_reports_near_root_only():
    o = np.zeros(3, np.float32)
    d = np.array([0, 0, 1], np.float32)
    c = np.zeros(3, np.float32)
    t = np.zeros(1, np.float32)
    assert orc.lib().orc_intersect_sphere(o.ctypes.data, d.ctypes.data, c.ctypes.data, 1.0,
                                          t.ctypes.data)
    assert t[0] == -1.0   # behind the origin: rejected by the [tmin, tmax] test


def test_triangle_matches_fp64_moller_trumbore():
    rng = np.random.default_rng(2)
    bad = 0
    for _ in range(3000):
        v = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = v.mean(0) + rng.normal(0, 0.5, 3)
        d = _unit(target - o).astype(np.float32)
        t, u, w = (np.zeros(1, np.float32) for _ in range(3))
        hit = orc.lib().orc_intersect_triangle(o.ctypes.data, d.ctypes.data, v[0].ctypes.data,
                                               v[1].ctypes.data, v[2].ctypes.data, t.ctypes.data,
                                               u.ctypes.data, w.ctypes.data)
        od, dd = o.astype(np.float64), d.astype(np.float64)
        vd = v.astype(np.float64)
        tr = orc.lib().orc_intersect_triangle_f64(od.ctypes.data, dd.ctypes.data,
                                                  vd[0].ctypes.data, vd[1].ctypes.data,
                                                  vd[2].ctypes.data)
        if not math.isnan(tr):
            if not hit:
                bad += 1
                continue
            assert abs(t[0] - tr) <= 1e-4 * max(1.0, abs(tr))
        elif hit:
            bad += 1
    assert bad <= 3   # edge-grazing rays where fp32 and fp64 legitimately disagree


def test_triangle_fan_is_watertight():
    """Rays through the shared vertex and edges of a fan hit at least one triangle."""
    n = 12
    verts = [np.zeros(3)] + [np.array([math.cos(2 * math.pi * k / n), math.sin(2 * math.pi * k / n),
                                       0.0]) for k in range(n)]
    tris = [(0, 1 + k, 1 + (k + 1) % n) for k in range(n)]
    rng = np.random.default_rng(3)
    for trial in range(400):
        k = trial % n
        a = verts[1 + k] * rng.uniform(0, 1) if trial % 3 else np.zeros(3)
        o = np.array([a[0], a[1], 2.0], np.float32)
        d = np.array([0, 0, -1], np.float32)
        hits = 0
        for (i, j, l) in tris:
            v0, v1, v2 = (np.asarray(verts[x], np.float32) for x in (i, j, l))
            t = np.zeros(1, np.float32)
            u = np.zeros(1, np.float32)
            w = np.zeros(1, np.float32)
            hits += orc.lib().orc_intersect_triangle(o.ctypes.data, d.ctypes.data, v0.ctypes.data,
                                                     v1.ctypes.data, v2.ctypes.data, t.ctypes.data,
                                                     u.ctypes.data, w.ctypes.data)
        assert hits >= 1


def test_cylinder_matches_fp64_quadratic():
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(3000):
        p0 = rng.uniform(-1, 1, 3).astype(np.float32)
        p1 = (p0 + rng.normal(0, 0.7, 3)).astype(np.float32)
        r = np.float32(rng.uniform(0.05, 0.4))
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        d = _unit((p0 + p1) / 2 - o + rng.normal(0, 0.3, 3)).astype(np.float32)
        ts = np.zeros(3, np.float32)
        ns = np.zeros(9, np.float32)
        valid = np.zeros(3, np.int32)
        orc.lib().orc_intersect_cylinder(o.ctypes.data, d.ctypes.data, p0.ctypes.data,
                                         p1.ctypes.data, float(r), 0, ts.ctypes.data,
                                         ns.ctypes.data, valid.ctypes.data)
        # fp64: |(q - p0) x s|^2 = r^2 |s|^2 with q = o + t d, 0 < (q-p0).s < |s|^2
        s = (p1 - p0).astype(np.float64)
        f = o.astype(np.float64) - p0
        dd = d.astype(np.float64)
        A = np.cross(dd, s)
        B = np.cross(f, s)
        a = A @ A
        b = 2 * (A @ B)
        c = B @ B - float(r) ** 2 * (s @ s)
        disc = b * b - 4 * a * c
        if a < 1e-9 or abs(disc) < 1e-6:
            continue
        roots = []
        if disc > 0:
            for tr in ((-b - math.sqrt(disc)) / (2 * a), (-b + math.sqrt(disc)) / (2 * a)):
                proj = (f + tr * dd) @ s
                if 1e-4 * (s @ s) < proj < (1 - 1e-4) * (s @ s):
                    roots.append(tr)
        got = sorted(float(ts[k]) for k in range(2) if valid[k])
        if len(roots) == len(got):
            for g, e in zip(got, sorted(roots)):
                assert abs(g - e) <= 2e-4 * max(1.0, abs(e))
            checked += 1
    assert checked > 2500


def _random_scene(seed):
    rng = np.random.default_rng(seed)
    sph = S.GeomDesc("sphere", rng.random((300, 3), dtype=np.float32),
                     radius=rng.uniform(0.005, 0.05, 300).astype(np.float32), surface_id=1)
    tv = rng.random((3 * 200, 3), dtype=np.float32)
    tri = S.GeomDesc("triangle", tv, surface_id=2)
    cp = rng.random((2 * 150, 3), dtype=np.float32)
    cyl = S.GeomDesc("cylinder", cp, radius=rng.uniform(0.003, 0.02, 150).astype(np.float32),
                     caps=True, surface_id=3)
    xfm = np.array([0, 1, 0, -1, 0, 0, 0, 0, 1, 0.2, 0.9, 0.1], np.float32)   # 90 deg about z
    sc = S.SceneDesc(geoms=[sph, tri, cyl], world_geoms=[0, 1],
                     instances=[S.InstanceDesc(geoms=[2], xfm=xfm, id=7)], world_id=4)
    return sc, rng


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bvh_closest_hit_equals_brute_force(seed):
    sc, rng = _random_scene(seed)
    os_ = orc.OracleScene(sc)
    for _ in range(600):
        o = rng.uniform(-0.5, 1.5, 3).astype(np.float32)
        d = _unit(rng.uniform(0, 1, 3) - o).astype(np.float32)
        a = os_.trace_closest(o, d)
        b = os_.trace_brute(o, d)
        assert (a is None) == (b is None)
        if a is not None:
            assert (a.t, a.entry, a.prim) == (b.t, b.entry, b.prim)


def _transform_array_scene(n=500, seed=3):
    """One instance whose transform array places a small sphere mesh n times
    (rotation, scale, translation per element) with per-element ids."""
    rng = np.random.default_rng(seed)
    v, idx = S.gen_sphere_mesh(80, seed=seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
                  2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                  2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)], 1)
    xf = np.concatenate([R * rng.uniform(0.01, 0.04, (n, 1)), rng.random((n, 3))], 1)
    inst = S.InstanceDesc(geoms=[0], xfm=xf.astype(np.float32),
                          ids=rng.permutation(n).astype(np.uint32) + 1000)
    sc = S.SceneDesc(geoms=[S.GeomDesc("triangle", v, idx, surface_id=3)], instances=[inst])
    return sc, rng


def test_transform_array_instances_and_entry_bvh():
    """Instance transform arrays (Instance.cpp:115-128): n entries with their
    own ids.  The oracle's entry-level BVH gives the brute force's closest
    hits, and frames equal those of a linear walk over the entries."""
    sc, rng = _transform_array_scene()
    os_ = orc.OracleScene(sc)
    assert os_.s.num_entries == 500 and os_.s.tlas
    for _ in range(400):
        o = rng.uniform(-0.5, 1.5, 3).astype(np.float32)
        d = _unit(rng.uniform(0, 1, 3) - o).astype(np.float32)
        a = os_.trace_closest(o, d)
        b = os_.trace_brute(o, d)
        assert (a is None) == (b is None)
        if a is not None:
            assert (a.t, a.entry, a.prim) == (b.t, b.entry, b.prim)
    cam = S.CameraDesc((0.5, 0.5, -1.0), (0, 0, 1), (0, 1, 0), math.radians(50), 1.0)
    with_bvh = os_.render(cam, 64, 64)
    with_bvh_ao = os_.render(cam, 64, 64, kind="ao", ambientSamples=2,
                             ambientOcclusionDistance=0.1)
    os_.s.tlas = None                     # linear walk (the structure stays allocated)
    linear = os_.render(cam, 64, 64)
    linear_ao = os_.render(cam, 64, 64, kind="ao", ambientSamples=2,
                           ambientOcclusionDistance=0.1)
    for a, b in ((with_bvh, linear), (with_bvh_ao, linear_ao)):
        for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
            np.testing.assert_array_equal(a[ch], b[ch], err_msg=ch)
    hit = with_bvh["depth"] < 1e29
    ids = set(np.unique(with_bvh["instanceId"][hit]).tolist())
    assert len(ids) > 50 and ids <= set(range(1000, 1500))


def test_tie_rule_lowest_entry_then_primitive():
    c = np.array([[0, 0, 0]] * 8, np.float32)
    c[3] = c[7] = (0.5, 0.5, 0.5)
    r = np.full(8, 0.1, np.float32)
    g = S.GeomDesc("sphere", c, radius=r)
    sc = S.SceneDesc(geoms=[g, g], world_geoms=[0, 1])   # the same spheres twice
    os_ = orc.OracleScene(sc)
    h = os_.trace_closest([0.5, 0.5, -1.0], [0, 0, 1])
    assert (h.entry, h.prim) == (0, 3)
    hb = os_.trace_brute([0.5, 0.5, -1.0], [0, 0, 1])
    assert (hb.entry, hb.prim) == (0, 3)


def test_camera_invariants():
    fovy = math.radians(45.0)
    cam = S.CameraDesc((0.5, -1.5, 1.2), (0.0, 2.0, -1.2), (0, 0, 1), fovy, 4 / 3)
    W, H = 1024, 768
    # pixel (0,0) is the bottom-left corner ray; (W/2, H/2) the centre ray
    _, dc = orc.camera_ray(cam, W, H, W / 2, H / 2)
    np.testing.assert_allclose(dc, _unit(cam.direction), atol=1e-6)
    _, d_bot = orc.camera_ray(cam, W, H, W / 2, 0)
    _, d_top = orc.camera_ray(cam, W, H, W / 2, H)
    ang = math.acos(np.clip(np.dot(d_bot.astype(np.float64), d_top), -1, 1))
    assert abs(ang - fovy) < 1e-5
    _, d_l = orc.camera_ray(cam, W, H, 0, H / 2)
    _, d_r = orc.camera_ray(cam, W, H, W, H / 2)
    ang_x = math.acos(np.clip(np.dot(d_l.astype(np.float64), d_r), -1, 1))
    assert abs(ang_x - 2 * math.atan(math.tan(fovy / 2) * 4 / 3)) < 1e-5
    # bottom rows look down (z decreases with the up vector (0,0,1))
    assert d_bot[2] < d_top[2]


def test_packing_known_answers():
    assert orc.pack_unorm4x8([0, 0, 0, 0]) == 0
    assert orc.pack_unorm4x8([1, 1, 1, 1]) == 0xFFFFFFFF
    assert orc.pack_unorm4x8([0.5, 0, 0, 0]) == 128          # round(127.5) = 128
    assert orc.pack_unorm4x8([-1, 2, 0.25, 1]) == (0xFF << 24) | (64 << 16) | (255 << 8)
    s = orc.linear_to_srgb([0.5, 0.001, 1.0, 0.3])
    assert abs(s[0] - (1.055 * 0.5 ** 0.41666 - 0.055)) < 2e-7
    assert abs(s[1] - 0.001 * 12.92) < 1e-9
    assert abs(s[2] - 1.0) < 2e-7 and s[3] == np.float32(0.3)


def _vol(data, lut, spacing=1.0 / 15):
    return S.VolumeDesc(data=data, spacing=(spacing,) * 3, color=lut[0], opacity=lut[1],
                        unit_distance=spacing)


def test_volume_constant_field_closed_form():
    """Constant field s, constant colour c and alpha a: after n samples
    C = c (1 - tau^n), O = 1 - tau^n with tau = (1-a)^(step/unitDistance)."""
    n = 16
    data = np.full((n, n, n), 0.5, np.float32)
    color = np.array([[0.2, 0.4, 0.8], [0.2, 0.4, 0.8]], np.float32)
    a = 0.03
    vol = _vol(data, (color, np.array([a, a], np.float32)))
    cam = S.CameraDesc((0.5, 0.5, -1.0), (0, 0, 1), (0, 1, 0), math.radians(10), 1.0)
    # 0.55-voxel step: the exit is not a multiple of the step (no knife edge)
    img, ns = orc.volume_image(vol, cam, 8, 8, inv_sampling_rate=1.1, early_stop=False)
    sp = np.float32(1.0 / 15)
    step = float(sp / np.float32(2.0)) * 1.1
    length = 15.0 * float(sp)            # axis-aligned centre ray through [0, 15*sp]
    nsamp = int(math.floor(length / step)) + 1
    tau = (1 - float(np.float32(a))) ** (step / float(sp))   # the LUT holds fp32 alpha
    px = img[4, 4]
    np.testing.assert_allclose(px[3], 1 - tau ** nsamp, rtol=1e-6)   # fp32 LUT rounding
    np.testing.assert_allclose(px[:3], color[0] * (1 - tau ** nsamp), rtol=1e-6)


def test_volume_linear_field_is_reproduced_exactly():
    n = 9
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    data = (0.03 * x + 0.05 * y + 0.02 * z + 0.1).astype(np.float32)
    v = orc.orc_volume()
    v.dims[:] = [n, n, n]
    v.origin[:] = [0.0, 0.0, 0.0]
    v.spacing[:] = [0.5, 0.25, 2.0]
    v.data = data.ctypes.data
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = np.array([rng.uniform(0, 4), rng.uniform(0, 2), rng.uniform(0, 16)])
        gx, gy, gz = p[0] / 0.5, p[1] / 0.25, p[2] / 2.0
        want = 0.03 * gx + 0.05 * gy + 0.02 * gz + 0.1
        got = orc.lib().orc_sample_field(orc.C.byref(v), p.ctypes.data)
        assert abs(got - want) < 1e-6


def test_tf_discretization_matches_interpolation():
    colors = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32)
    op = np.array([0.0, 0.05], np.float32)
    lut = orc.tf_discretize(colors, op)
    p = np.arange(256) / 255.0
    np.testing.assert_allclose(lut[:, 3], 0.05 * p, atol=1e-7)
    for k in range(3):
        np.testing.assert_allclose(lut[:, k], np.interp(p, [0, 0.5, 1], colors[:, k]), atol=1e-6)


def test_value_range_clamp_follows_position():
    """position() (gpu_math.h:185-189) is glm::clamp(s, lo, hi) = min(max(s,
    lo), hi), then (s - lo) / (hi - lo), read at texel u * 256 - 0.5: values
    outside a range read the end texels, every value of a reversed range
    (lo > hi) reads the last texel, and an empty range reads texel 0."""
    from visrtx_amd import scenes as S
    colors = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32)
    op = np.array([0.1, 0.9], np.float32)
    data = np.zeros((2, 2, 2), np.float32)
    for vr, cases in (((0.2, 0.8), {-5.0: 0, 0.1: 0, 0.2: 0, 0.8: 255, 3.0: 255}),
                      ((0.8, 0.2), {-5.0: 255, 0.1: 255, 0.5: 255, 0.9: 255}),
                      ((0.5, 0.5), {-5.0: 0, 0.5: 0, 7.0: 0})):
        vol = S.VolumeDesc(data, color=colors, opacity=op, value_range=vr)
        lut = orc.tf_discretize(colors, op, vr)
        for s, texel in cases.items():
            np.testing.assert_allclose(orc.classify(vol, s), lut[texel], rtol=0, atol=1e-12,
                                       err_msg=f"range {vr}, s = {s}")
    vol = S.VolumeDesc(data, color=colors, opacity=op, value_range=(0.2, 0.8))
    lut = orc.tf_discretize(colors, op, (0.2, 0.8))
    # inside: texel coordinate (s - lo) / (hi - lo) * 256 - 0.5, linear (the
    # range is stored as fp32)
    lo, hi = float(np.float32(0.2)), float(np.float32(0.8))
    x = (0.5 - lo) / (hi - lo) * 256 - 0.5
    i = int(np.floor(x))
    a, b = lut[i].astype(np.float64), lut[i + 1].astype(np.float64)
    want = a + (b - a) * (x - i)
    np.testing.assert_allclose(orc.classify(vol, 0.5), want, rtol=0, atol=1e-9)


@pytest.mark.parametrize("seed", [3, 4])
def test_any_hit_equals_closest_hit_within_range(seed):
    """Opaque geometry: the AO any-hit query is 'a closest hit exists in [tmin, tmax]'."""
    sc, rng = _random_scene(seed)
    os_ = orc.OracleScene(sc)
    agree = 0
    for _ in range(600):
        o = rng.uniform(-0.5, 1.5, 3).astype(np.float32)
        d = _unit(rng.uniform(0, 1, 3) - o).astype(np.float32)
        tmax = float(rng.uniform(0.05, 2.0))
        h = os_.trace_closest(o, d, 0.0, tmax)
        assert os_.trace_any(o, d, 0.0, tmax) == (h is not None)
        agree += h is not None
    assert agree > 50


def _cube_mesh():
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    idx = []
    for a, b, c, d in quads:
        idx += [(a, b, c), (a, c, d)]
    return v, np.array(idx, np.uint32)


def test_ao_closed_forms_unoccluded_and_enclosed():
    """AO = 1 with nothing above a floor (colour = base * radiance * ambientColor);
    AO = 0 inside a closed, index-shared cube (every AO ray hits a wall)."""
    floor_v = np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float32)
    floor_i = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    sc = S.SceneDesc(geoms=[S.GeomDesc("triangle", floor_v, floor_i, color=(0.8, 0.6, 0.4))],
                     world_geoms=[0])
    cam = S.CameraDesc((0.0, 2.0, -3.0), (0.0, -2.0, 3.0), (0, 1, 0), math.radians(40), 1.0)
    o = orc.OracleScene(sc).render(cam, 24, 24, "ao", pixelSamples=1, ambientSamples=8,
                                   ambientRadiance=0.5, ambientColor=(1.0, 0.5, 0.25))
    hit = o["depth"] < 1e29
    assert hit.mean() > 0.3
    want = np.array([0.8 * 0.5 * 1.0, 0.6 * 0.5 * 0.5, 0.4 * 0.5 * 0.25], np.float32)
    np.testing.assert_allclose(o["color"][hit][:, :3], np.broadcast_to(want, (hit.sum(), 3)),
                               rtol=2e-6)
    np.testing.assert_allclose(o["color"][hit][:, 3], 1.0, rtol=1e-6)

    v, idx = _cube_mesh()
    sc = S.SceneDesc(geoms=[S.GeomDesc("triangle", v, idx)], world_geoms=[0])
    cam = S.CameraDesc((0.5, 0.5, 0.5), (0.3, 0.2, 1.0), (0, 1, 0), math.radians(60), 1.0)
    o = orc.OracleScene(sc).render(cam, 24, 24, "ao", pixelSamples=2, ambientSamples=16)
    assert (o["depth"] < 1e29).all()
    assert (o["color"][..., :3] == 0.0).all()
    np.testing.assert_allclose(o["color"][..., 3], 1.0, rtol=1e-6)


def test_diffvol_forward_equals_volume_march():
    """The FP64 march of the differentiable path (explicit rays, explicit LUT)
    equals the FP64 march of the renderer path on the same field and rays."""
    n = 12
    data = S.gen_noise_volume(n, seed=9)
    vol = _vol(data, (S.COLORMAP5, np.array([0.0, 0.4], np.float32)), spacing=1.0 / (n - 1))
    W = H = 10
    cam = S.CameraDesc((1.9, 1.5, -1.0), S.look_at((1.9, 1.5, -1.0), (0.5, 0.5, 0.5)), (0, 1, 0),
                       math.radians(35), 1.0)
    img, _ = orc.volume_image(vol, cam, W, H, inv_sampling_rate=1.0, early_stop=True)
    rays = [orc.camera_ray(cam, W, H, x, y) for y in range(H) for x in range(W)]
    org = np.array([r[0] for r in rays], np.float32)
    dirs = np.array([r[1] for r in rays], np.float32)
    sp = np.float32(vol.spacing[0])
    step = float(sp / np.float32(2.0))
    lut = orc.tf_discretize(vol.color, vol.opacity, vol.value_range)
    got, _ = orc.diffvol_forward((n, n, n), vol.origin, vol.spacing, data, lut, vol.value_range,
                                 vol.unit_distance, step, org, dirs, early_stop=True)
    # the renderer path derives its step in fp32 (min(spacing)/2 * 1/rate): differences
    # stay at the level of that one rounding
    np.testing.assert_allclose(got, img.reshape(-1, 4), rtol=1e-7, atol=1e-9)
    assert img[..., 3].max() > 0.3


def test_quad_split_matches_reference_rule():
    """Quad.cpp:124-145: indexed quad (x, y, z, w) -> triangles (x, y, w) and
    (z, w, y); soup quad q -> 4q + (0, 1, 3) and 4q + (2, 3, 1)."""
    q = np.array([[10, 11, 12, 13], [7, 3, 5, 1]], np.uint32)
    t = S.quads_to_triangles(q, 20)
    assert t.tolist() == [[10, 11, 13], [12, 13, 11], [7, 3, 1], [5, 1, 3]]
    ts = S.quads_to_triangles(None, 8)
    assert ts.tolist() == [[0, 1, 3], [2, 3, 1], [4, 5, 7], [6, 7, 5]]


def _cone(o, d, p0, p1, ra, rb):
    ts = np.zeros(2, np.float32)
    ns = np.zeros(6, np.float32)
    valid = np.zeros(2, np.int32)
    keep = [np.asarray(v, np.float32) for v in (o, d, p0, p1)]
    r = orc.lib().orc_intersect_cone(*[k.ctypes.data for k in keep], ra, rb,
                                     ts.ctypes.data, ns.ctypes.data, valid.ctypes.data)
    return r, ts, ns.reshape(2, 3), valid


def test_cone_intersection_closed_forms():
    """Intersectors_ptx.cu:192-246 restated: equal radii reduce to a cylinder,
    a ray down the axis meets the near cap, a pointed cone's side is hit where
    its radius equals the ray's distance from the axis, and misses miss."""
    # ra == rb: side hit at distance 0.5 from the x axis
    r, ts, ns, v = _cone((0, 0, -5), (0, 0, 1), (-1, 0, 0), (1, 0, 0), 0.5, 0.5)
    assert v[1] and not v[0] and ts[1] == pytest.approx(4.5, rel=1e-6)
    n = ns[1] / np.linalg.norm(ns[1])
    np.testing.assert_allclose(n, [0, 0, -1], atol=1e-6)
    # down the axis from beyond p0: the p0 cap, normal -axis
    r, ts, ns, v = _cone((-3, 0, 0), (1, 0, 0), (-1, 0, 0), (1, 0, 0), 0.5, 0.25)
    assert v[0] and ts[0] == pytest.approx(2.0, rel=1e-6)
    np.testing.assert_allclose(ns[0] / np.linalg.norm(ns[0]), [-1, 0, 0], atol=1e-6)
    # from beyond p1: the p1 cap
    r, ts, ns, v = _cone((4, 0.1, 0), (-1, 0, 0), (-1, 0, 0), (1, 0, 0), 0.5, 0.25)
    assert v[0] and ts[0] == pytest.approx(3.0, rel=1e-6)
    # pointed cone along z (radius 1 at z=0, 0 at z=2): at z = 1 the radius is 0.5
    r, ts, ns, v = _cone((5, 0, 1), (-1, 0, 0), (0, 0, 0), (0, 0, 2), 1.0, 0.0)
    assert v[1] and ts[1] == pytest.approx(4.5, rel=1e-6)
    n = ns[1] / np.linalg.norm(ns[1])
    np.testing.assert_allclose(n, np.array([2.0, 0.0, 1.0]) / math.sqrt(5.0), atol=1e-6)
    # beside the pointed cone: miss
    r, ts, ns, v = _cone((5, 0.5, 1.0), (-1, 0, 0), (0, 0, 0), (0, 0, 2), 0.2, 0.0)
    assert not v.any()


def test_cone_bounds_use_the_larger_end_radius():
    """Cone.cpp:98-105: box = min/max of the end points -/+ max(r0, r1)."""
    pos = np.array([[0, 0, 0], [1, 2, 3]], np.float32)
    box = orc.bounds(orc.ORC_CONE, pos, radius=np.array([0.5, 0.25], np.float32))
    np.testing.assert_allclose(box.reshape(-1), [-0.5, -0.5, -0.5, 1.5, 2.5, 3.5])


def _round_cone(o, d, pa, pb, ra, rb):
    t = np.zeros(1, np.float32)
    n = np.zeros(3, np.float32)
    keep = [np.asarray(v, np.float32) for v in (o, d, pa, pb)]
    r = orc.lib().orc_intersect_round_cone(*[k.ctypes.data for k in keep], ra, rb,
                                           t.ctypes.data, n.ctypes.data)
    return r, float(t[0]), n


def _swept_sphere_entry(o, d, pa, pb, ra, rb, steps=20001):
    """FP64 brute force: the first entry of the ray into any sphere of the
    sweep c(s) = pa + s (pb - pa), r(s) = ra + s (rb - ra), s in [0, 1]."""
    o, d, pa, pb = (np.asarray(v, np.float64) for v in (o, d, pa, pb))
    s = np.linspace(0.0, 1.0, steps)
    c = pa[None] + s[:, None] * (pb - pa)[None]
    r = ra + s * (rb - ra)
    oc = o[None] - c
    b = oc @ d
    h = b * b - (np.sum(oc * oc, 1) - r * r)
    ok = h >= 0
    if not ok.any():
        return None
    return float(np.min(-b[ok] - np.sqrt(h[ok])))


def test_round_linear_curve_against_a_swept_sphere():
    """The closed form (tangent cone between the end spheres, else the nearer
    end sphere) against an FP64 brute force over 20,001 spheres of the sweep,
    on random rays and segments; hit/miss agree away from grazing rays."""
    rng = np.random.default_rng(12)
    checked = disagree = 0
    for _ in range(300):
        pa = rng.uniform(-0.5, 0.5, 3)
        pb = pa + rng.normal(0, 0.6, 3)
        ra, rb = rng.uniform(0.05, 0.4, 2)
        o = rng.uniform(-3, 3, 3)
        tgt = (pa + pb) / 2 + rng.normal(0, 0.4, 3)
        d = tgt - o
        d /= np.linalg.norm(d)
        hit, t, n = _round_cone(o, d, pa, pb, ra, rb)
        tb = _swept_sphere_entry(o, d, pa, pb, ra, rb)
        if (tb is None) != (not hit):
            disagree += 1           # only a ray grazing the surface may differ
            continue
        if tb is None:
            continue
        assert t == pytest.approx(tb, rel=2e-4, abs=2e-4)
        checked += 1
    assert checked > 80 and disagree <= 2


def test_round_linear_curve_special_cases():
    # equal radii: a capsule; side hit at distance r from the axis
    hit, t, n = _round_cone((0, 0, -5), (0, 0, 1), (-1, 0, 0), (1, 0, 0), 0.5, 0.5)
    assert hit and t == pytest.approx(4.5, rel=1e-6)
    np.testing.assert_allclose(n / np.linalg.norm(n), [0, 0, -1], atol=1e-6)
    # down the axis: the end sphere of p0 (rounded end, not a flat cap)
    hit, t, n = _round_cone((-3, 0, 0), (1, 0, 0), (-1, 0, 0), (1, 0, 0), 0.5, 0.25)
    assert hit and t == pytest.approx(1.5, rel=1e-6)
    # the small sphere inside the large one: the large sphere alone
    hit, t, n = _round_cone((0, 0, -5), (0, 0, 1), (0, 0, 0), (0.1, 0, 0), 1.0, 0.2)
    assert hit and t == pytest.approx(4.0, rel=1e-6)
    # a miss
    hit, t, n = _round_cone((0, 3, -5), (0, 0, 1), (-1, 0, 0), (1, 0, 0), 0.5, 0.5)
    assert not hit


def test_curve_bounds_cover_both_end_spheres():
    pos = np.array([[0, 0, 0], [1, 2, 3], [4, 4, 4]], np.float32)
    box = orc.bounds(orc.ORC_CURVE, pos, index=np.array([0, 1], np.uint32),
                     radius=np.array([0.5, 0.25, 1.0], np.float32))
    np.testing.assert_allclose(box, [[-0.5, -0.5, -0.5, 1.25, 2.25, 3.25],
                                     [0.75, 1.75, 2.75, 5.0, 5.0, 5.0]])
