"""GPU results against the committed golden fixtures (tests/golden/fixtures.npz):
LBVH arrays bit-exact, frame ids exact with depth and colour within 1e-4
relative, volume colour within 1e-4 of the FP64 values on pixels the oracle
does not flag as fp32-ambiguous."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as MF  # noqa: E402
from tests.test_gpu_lbvh import gpu_build  # noqa: E402
from tests.test_gpu_render import render_gpu  # noqa: E402

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(HERE, "golden", "fixtures.npz"))


def test_lbvh_matches_fixture(require_gpu):
    from visrtx_amd import _native as N
    g = gpu_build(N.VRX_SPHERE, FIX["lbvh_centers"], radius=FIX["lbvh_radii"])
    np.testing.assert_array_equal(g["codes"], FIX["lbvh_sorted_codes"])
    np.testing.assert_array_equal(g["ids"], FIX["lbvh_order"])
    np.testing.assert_array_equal(g["child"], FIX["lbvh_child"])
    np.testing.assert_array_equal(g["boxes"], FIX["lbvh_node_boxes"])
    np.testing.assert_array_equal(g["root_box"], FIX["lbvh_root_box"])


@pytest.mark.parametrize("name", ["spheres", "instanced_triangles", "cylinders_caps"])
def test_frames_match_fixture(require_gpu, name):
    surf, _ = MF.fixture_scenes()
    sc, cam = surf[name]
    g = render_gpu(sc, cam, MF.W, MF.H, background=(0.1, 0.2, 0.3, 1.0))
    assert not g["errors"], g["errors"]
    for ch in ("primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(g[ch], FIX[f"{name}_{ch}"], err_msg=ch)
    np.testing.assert_allclose(g["depth"], FIX[f"{name}_depth"], rtol=1e-4, atol=0)
    np.testing.assert_allclose(g["color"], FIX[f"{name}_color"], rtol=1e-4, atol=1e-4)


def test_volume_matches_fixture(require_gpu):
    _, (sc_v, vol, cam_v) = MF.fixture_scenes()
    g = render_gpu(sc_v, cam_v, MF.W, MF.H, volumeSamplingRate=1.0)
    assert not g["errors"], g["errors"]
    ok = FIX["volume_frame_ambiguous"] == 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(g["color"][ok], FIX["volume_frame_color"][ok], rtol=0, atol=1e-4)
    for ch in ("primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(g[ch], FIX[f"volume_frame_{ch}"], err_msg=ch)
    np.testing.assert_allclose(g["depth"], FIX["volume_frame_depth"], rtol=1e-5)
