"""Traversal at both BVH node widths (renderer nodeWidth = 2 and 4; the
device default is 4) vs the oracle: the closest hit is the lexicographic
minimum over all intersections, so the node width must not change a single
id, depth or colour bit."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S
from tests.test_gpu_render import (_cylinder_scene, _instanced_scene, compare, render_gpu)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [2, 4])
def test_wide_spheres(require_gpu, width):
    c, r = S.gen_spheres(50_000, seed=31, rmin=2e-3, rmax=1e-2)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", c, radius=r, surface_id=2)], world_geoms=[0],
                     world_id=1)
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), math.radians(45.0), 1.5)
    g = render_gpu(sc, cam, 240, 160, nodeWidth=width)
    o = orc.OracleScene(sc).render(cam, 240, 160)
    compare(g, o)


@pytest.mark.parametrize("width", [2, 4])
def test_wide_instanced_triangles(require_gpu, width):
    sc, cam = _instanced_scene()
    g = render_gpu(sc, cam, 200, 200, nodeWidth=width)
    o = orc.OracleScene(sc).render(cam, 200, 200)
    compare(g, o)
    assert len(np.unique(g["instanceId"])) > 3


@pytest.mark.parametrize("caps", [False, True])
def test_wide_cylinders(require_gpu, caps):
    sc, cam = _cylinder_scene(caps)
    g = render_gpu(sc, cam, 160, 120, nodeWidth=2)
    o = orc.OracleScene(sc).render(cam, 160, 120)
    compare(g, o)


def test_wide_ao_equals_binary(require_gpu):
    """AO (any-hit occlusion rays) with 4-wide nodes equals the BVH2 result."""
    sc, cam = _instanced_scene()
    kw = dict(pixelSamples=2, ambientSamples=4, ambientOcclusionDistance=0.5)
    a = render_gpu(sc, cam, 96, 96, "ao", nodeWidth=2, **kw)
    b = render_gpu(sc, cam, 96, 96, "ao", nodeWidth=4, **kw)
    for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(a[ch], b[ch], err_msg=ch)
