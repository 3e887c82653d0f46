"""Pins the oracle's raycast colour path (Raycast_ptx.cu:60-179) and the
frame's tonemapped accumulation (gpu_util.h:323-443) with closed forms.

Each expected value is computed here in FP64 from the scene description
alone: the camera basis of camera/Perspective.cpp:40-72, the pixel-corner
primary ray of gpu/cameraCreateRay.h:74-81 (row 0 at the bottom), an
analytic ray/plane intersection, and the raycast layer loop:

    lighting = |dot(dir, Ns)| * ambientColor
    color    = base * lighting,  opacity = material opacity
    color   *= opacity                                  (Raycast_ptx.cu:136)
    out     += color * (1 - outOpacity)    accumulateValue, gpu_util.h:130-134
    outOp   += opacity * (1 - outOpacity)
    miss:    out += background.rgb * (1 - outOpacity)
             outOp += background.a * (1 - outOpacity)

Ns is the barycentric blend of the vertex normals, normalised and turned
into Ng's hemisphere (populateHit.h:206-264); the turn is invisible to the
raycast colour (|dot|), so the blend and the normalisation are what a
vertex-normal frame pins.  The accumulation check uses the AO renderer,
whose samples are averaged in tonemapped space:
    pixel = inverseTonemap(sum_i tonemap(c_i) / N),
    tonemap(c) = c / (1 + max(0, max c)),  inverseTonemap(c) = c / max(1e-12, 1 - max c).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S

W, H = 40, 30


def camera_rays(cam, w, h, ox=0.0, oy=0.0):
    """FP64 primary rays (origin, unit directions (h, w, 3)) through the
    pixel corners (x + ox, y + oy) / size, Perspective.cpp:40-72."""
    pos = np.asarray(cam.position, np.float64)
    d = np.asarray(cam.direction, np.float64)
    d = d / np.linalg.norm(d)
    up = np.asarray(cam.up, np.float64)
    up = up / np.linalg.norm(up)
    ih = 2.0 * math.tan(0.5 * cam.fovy)
    iw = ih * cam.aspect
    du = np.cross(d, up)
    du = du / np.linalg.norm(du) * iw
    dv = np.cross(du, d)
    dv = dv / np.linalg.norm(dv) * ih
    d00 = d - 0.5 * du - 0.5 * dv
    xs = (np.arange(w, dtype=np.float64) + ox) / w
    ys = (np.arange(h, dtype=np.float64) + oy) / h
    dirs = d00[None, None, :] + xs[None, :, None] * du[None, None, :] + \
        ys[:, None, None] * dv[None, None, :]
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    return pos, dirs


def plane_t(pos, dirs, p0, n):
    """Distance along each ray to the plane through p0 with normal n."""
    n = np.asarray(n, np.float64)
    return (np.dot(np.asarray(p0, np.float64) - pos, n)) / (dirs @ n)


def square(z, half, color, opacity=1.0, sid=1, normals=None, flip=False):
    """Square |x|,|y| <= half in the plane z, as two triangles (normal +-z)."""
    v = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]],
                 np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]] if not flip else [[0, 2, 1], [0, 3, 2]], np.uint32)
    return S.GeomDesc("triangle", v, idx, color=color, opacity=opacity, surface_id=sid,
                      vertex_normal=normals)


def _scene(geoms):
    return S.SceneDesc(geoms=geoms, world_geoms=list(range(len(geoms))), world_id=0)


CAM = S.CameraDesc((0.1, -0.2, -3.0), (0.05, 0.1, 1.0), (0.0, 1.0, 0.0), math.radians(40.0),
                   W / H)


@pytest.mark.parametrize("flip", [False, True])
def test_opaque_square_is_base_times_cosine_times_ambient(flip):
    """One opaque square filling the view: rgb = base * |cos| * ambientColor
    per pixel (the cosine of each pixel's analytic ray with the plane
    normal), alpha 1, depth = the analytic plane distance; the square seen
    from its back side gives the same colour (|dot|, Ng flipped)."""
    base, amb = (0.9, 0.5, 0.2), (1.0, 0.75, 0.5)
    sc = _scene([square(0.5, 20.0, base, flip=flip)])
    o = orc.OracleScene(sc).render(CAM, W, H, ambientColor=amb)
    pos, dirs = camera_rays(CAM, W, H)
    cos = np.abs(dirs[..., 2])
    want = cos[..., None] * (np.array(base) * np.array(amb))[None, None, :]
    np.testing.assert_allclose(o["color"][..., :3], want, rtol=2e-6, atol=0)
    np.testing.assert_allclose(o["color"][..., 3], 1.0, rtol=0, atol=1e-6)
    t = plane_t(pos, dirs, (0, 0, 0.5), (0, 0, 1))
    np.testing.assert_allclose(o["depth"], t, rtol=2e-6)
    assert np.ptp(cos) > 0.05     # the cosine really varies over the frame


def test_tilted_square_cosine():
    """A square tilted by 50 degrees about x: the cosine is |dir . n| with n
    the tilted normal, so a transposed or dropped term in the shading dot
    product changes every pixel."""
    a = math.radians(50.0)
    R = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    g = square(0.0, 20.0, (0.6, 0.7, 0.8))
    g.position = (g.position.astype(np.float64) @ R.T + [0, 0, 1.0]).astype(np.float32)
    sc = _scene([g])
    o = orc.OracleScene(sc).render(CAM, W, H)
    pos, dirs = camera_rays(CAM, W, H)
    n = R @ np.array([0, 0, 1.0])
    cos = np.abs(dirs @ n)
    np.testing.assert_allclose(o["color"][..., :3], cos[..., None] * [0.6, 0.7, 0.8],
                               rtol=1e-5, atol=0)
    np.testing.assert_allclose(o["depth"], plane_t(pos, dirs, (0, 0, 1.0), n), rtol=1e-5)


def test_vertex_normals_are_blended_and_normalised():
    """Ns = normalize(b0 n0 + b1 n1 + b2 n2) with the hit's barycentrics
    (populateHit.h:224-233): one triangle in z = 0.5 with three different
    vertex normals, one of them facing away from Ng; colour = base *
    |dir . Ns| per pixel, barycentrics solved here in FP64."""
    v = np.array([[-3.0, -3.0, 0.5], [3.0, -3.0, 0.5], [0.0, 3.0, 0.5]], np.float32)
    nrm = np.array([[0.0, 0.0, 1.0], [0.8, 0.0, 0.6], [0.0, 0.6, -0.8]], np.float32)
    g = S.GeomDesc("triangle", v, np.array([[0, 1, 2]], np.uint32), color=(1.0, 1.0, 1.0),
                   vertex_normal=nrm)
    sc = _scene([g])
    o = orc.OracleScene(sc).render(CAM, W, H)
    pos, dirs = camera_rays(CAM, W, H)
    t = plane_t(pos, dirs, (0, 0, 0.5), (0, 0, 1))
    p = pos + t[..., None] * dirs
    v64 = v.astype(np.float64)
    e1, e2 = v64[1] - v64[0], v64[2] - v64[0]
    M = np.array([[e1[0], e2[0]], [e1[1], e2[1]]])
    uv = np.linalg.solve(M, (p[..., :2] - v64[0, :2]).reshape(-1, 2).T).T.reshape(H, W, 2)
    b1, b2 = uv[..., 0], uv[..., 1]
    b0 = 1.0 - b1 - b2
    inside = (b0 > 1e-4) & (b1 > 1e-4) & (b2 > 1e-4)
    assert inside.mean() > 0.2
    ns = b0[..., None] * nrm[0] + b1[..., None] * nrm[1] + b2[..., None] * nrm[2]
    ns /= np.linalg.norm(ns, axis=-1, keepdims=True)
    want = np.abs(np.sum(dirs * ns, axis=-1))
    np.testing.assert_allclose(o["color"][..., 0][inside], want[inside], rtol=1e-5, atol=1e-6)
    # flat shading would be |dir_z|: the blend must change the picture
    assert np.abs(want[inside] - np.abs(dirs[..., 2][inside])).max() > 0.1
    np.testing.assert_array_equal(o["primitiveId"][inside], 0)
    assert (o["depth"][~inside & (np.minimum(np.minimum(b0, b1), b2) < -1e-3)] > 1e29).all()


@pytest.mark.parametrize("a1,a2", [(0.4, 0.7), (0.25, 0.25), (0.9, 0.5)])
def test_translucent_layers_composite_front_to_back(a1, a2):
    """Two translucent squares (alphaMode blend) in front of the background:
    rgb = L c1 a1 + L c2 a2 (1 - a1) + bg (1 - a1)(1 - a2),
    alpha = a1 + a2 (1 - a1) + bgA (1 - a1)(1 - a2),
    with L = |cos| * ambientColor (Raycast_ptx.cu:132-142, 159-165); the
    first hit's depth and ids are the front square's."""
    c1, c2 = np.array([0.9, 0.1, 0.3]), np.array([0.2, 0.8, 0.6])
    bg = np.array([0.05, 0.1, 0.9, 0.6])
    amb = np.array([1.0, 0.9, 0.8])
    sc = _scene([square(0.5, 20.0, tuple(c1), opacity=a1, sid=11),
                 square(1.5, 20.0, tuple(c2), opacity=a2, sid=22)])
    o = orc.OracleScene(sc).render(CAM, W, H, background=tuple(bg), ambientColor=tuple(amb))
    pos, dirs = camera_rays(CAM, W, H)
    L = np.abs(dirs[..., 2])[..., None] * amb
    rgb = L * c1 * a1 + L * c2 * a2 * (1 - a1) + bg[:3] * (1 - a1) * (1 - a2)
    alpha = a1 + a2 * (1 - a1) + bg[3] * (1 - a1) * (1 - a2)
    np.testing.assert_allclose(o["color"][..., :3], rgb, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(o["color"][..., 3], alpha, rtol=1e-5)
    np.testing.assert_array_equal(o["objectId"], 11)
    np.testing.assert_allclose(o["depth"], plane_t(pos, dirs, (0, 0, 0.5), (0, 0, 1)), rtol=2e-6)


def test_opaque_layer_ends_the_loop():
    """An opaque square in front of a translucent one hides it completely
    (outputOpacity reaches 1 >= 0.99 after one layer)."""
    sc = _scene([square(0.5, 20.0, (0.3, 0.6, 0.9), opacity=1.0, sid=1),
                 square(0.2, 20.0, (1.0, 0.0, 0.0), opacity=0.3, sid=2)])
    o = orc.OracleScene(sc).render(CAM, W, H, background=(1.0, 1.0, 1.0, 1.0))
    pos, dirs = camera_rays(CAM, W, H)
    L = np.abs(dirs[..., 2])[..., None]
    c = 0.3 * L * np.array([1.0, 0.0, 0.0]) + 0.7 * L * np.array([0.3, 0.6, 0.9])
    np.testing.assert_allclose(o["color"][..., :3], c, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(o["color"][..., 3], 1.0, atol=1e-6)
    np.testing.assert_array_equal(o["objectId"], 2)


def test_miss_writes_background_and_reference_miss_ids():
    """Nothing in view: colour = background (rgb and alpha, no premultiply),
    depth 1e30 is kept from the loop, primitiveId 0 and objectId /
    instanceId 0xFFFFFFFF (Raycast_ptx.cu:73-77, 150-157)."""
    sc = _scene([square(-5.0, 0.1, (1, 1, 1))])     # behind the camera
    bg = (0.2, 0.4, 0.6, 0.5)
    o = orc.OracleScene(sc).render(CAM, W, H, background=bg)
    np.testing.assert_allclose(o["color"], np.broadcast_to(np.array(bg, np.float32), (H, W, 4)),
                               rtol=1e-6)
    assert (o["depth"] == np.float32(1e30)).all()
    assert (o["primitiveId"] == 0).all()
    assert (o["objectId"] == 0xFFFFFFFF).all() and (o["instanceId"] == 0xFFFFFFFF).all()


def _tm(c):
    return c / (1.0 + max(0.0, float(np.max(c))))


def _itm(c):
    return c / max(1e-12, 1.0 - float(np.max(c)))


@pytest.mark.parametrize("spp", [1, 3, 8])
def test_accumulated_samples_of_a_constant_image_equal_one_sample(spp):
    """An unoccluded floor under the AO renderer gives every sample the same
    HDR colour c = base * ambientRadiance * ambientColor (max > 1, so the
    tonemap is far from the identity); inverseTonemap(N tonemap(c) / N)
    must give c back for any N."""
    floor_v = np.array([[-9, 0, -9], [9, 0, -9], [9, 0, 9], [-9, 0, 9]], np.float32)
    g = S.GeomDesc("triangle", floor_v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32),
                   color=(0.8, 0.6, 0.4))
    sc = _scene([g])
    cam = S.CameraDesc((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), math.radians(40), 1.0)
    o = orc.OracleScene(sc).render(cam, 24, 24, "ao", pixelSamples=spp, ambientSamples=2,
                                   ambientRadiance=3.0, ambientColor=(1.0, 0.5, 2.0))
    want = np.array([0.8 * 3.0, 0.6 * 3.0 * 0.5, 0.4 * 3.0 * 2.0])
    np.testing.assert_allclose(o["color"][..., :3], np.broadcast_to(want, (24, 24, 3)), rtol=3e-6)
    np.testing.assert_allclose(o["color"][..., 3], 1.0, rtol=1e-6)


def test_samples_are_averaged_in_tonemapped_space():
    """A pixel whose jittered samples land on two floors of HDR colours c1
    and c2 shows inverseTonemap((k tonemap(c1) + (N-k) tonemap(c2)) / N) for
    some k in 0..N (gpu_util.h:393-443); a plain average of the colours
    would match no such k.  Pixels along the seam must mix (0 < k < N)."""
    N = 8
    c1, c2 = np.array([2.4, 0.3, 0.3]), np.array([0.2, 0.2, 1.8])
    left = np.array([[-9, 0, -9], [0, 0, -9], [0, 0, 9], [-9, 0, 9]], np.float32)
    right = left + np.array([9, 0, 0], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    sc = _scene([S.GeomDesc("triangle", left, idx, color=tuple(c1 / 3.0), surface_id=1),
                 S.GeomDesc("triangle", right, idx, color=tuple(c2 / 3.0), surface_id=2)])
    cam = S.CameraDesc((0.013, 3.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), math.radians(40),
                       1.0)
    o = orc.OracleScene(sc).render(cam, 24, 24, "ao", pixelSamples=N, ambientSamples=1,
                                   ambientRadiance=3.0)
    cand = np.array([_itm((k * _tm(c1) + (N - k) * _tm(c2)) / N) for k in range(N + 1)])
    got = o["color"][..., :3].reshape(-1, 3).astype(np.float64)
    err = np.abs(got[:, None, :] - cand[None, :, :]).max(axis=-1) / np.abs(cand).max(axis=-1)
    k = err.argmin(axis=1)
    assert err.min(axis=1).max() < 1e-5
    mixed = (k > 0) & (k < N)
    assert mixed.sum() >= 10
    # a plain average of the colours is far from what mixed pixels show
    for kk in np.unique(k[mixed]):
        lin = (kk * c1 + (N - kk) * c2) / N
        assert np.abs(lin - cand[kk]).max() > 0.05
