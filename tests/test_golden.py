"""The committed golden fixtures (tests/golden/fixtures.npz) against the
oracle and the native generators on the CPU: the generators still produce the
stored inputs, and the oracle still produces the stored outputs bit for bit.
A change to either shows up here before any GPU comparison is affected."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from oracle import oracle as orc  # noqa: E402
from visrtx_amd import scenes as S  # noqa: E402

import make_fixtures as MF  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "fixtures.npz"))
CHANNELS = ("color", "depth", "primitiveId", "objectId", "instanceId")


def test_generators_reproduce_fixture_inputs():
    c, r = S.gen_spheres(1000, seed=40)
    np.testing.assert_array_equal(c, FIX["lbvh_centers"])
    np.testing.assert_array_equal(r, FIX["lbvh_radii"])
    np.testing.assert_array_equal(S.gen_noise_volume(20, seed=46), FIX["volume_data"])


def test_oracle_lbvh_matches_fixture():
    boxes = orc.bounds(orc.ORC_SPHERE, FIX["lbvh_centers"], radius=FIX["lbvh_radii"])
    codes, _ = orc.morton(boxes)
    sc, order = orc.sort(codes)
    child, _ = orc.karras(sc)
    nb, root = orc.refit(boxes, order, child)
    np.testing.assert_array_equal(sc, FIX["lbvh_sorted_codes"])
    np.testing.assert_array_equal(order, FIX["lbvh_order"])
    np.testing.assert_array_equal(child, FIX["lbvh_child"])
    np.testing.assert_array_equal(nb, FIX["lbvh_node_boxes"])
    np.testing.assert_array_equal(root, FIX["lbvh_root_box"])


@pytest.mark.parametrize("name", ["spheres", "instanced_triangles", "cylinders_caps"])
def test_oracle_frames_match_fixture(name):
    surf, _ = MF.fixture_scenes()
    sc, cam = surf[name]
    o = orc.OracleScene(sc).render(cam, MF.W, MF.H, background=(0.1, 0.2, 0.3, 1.0))
    for ch in CHANNELS:
        np.testing.assert_array_equal(o[ch], FIX[f"{name}_{ch}"], err_msg=ch)
    assert (FIX[f"{name}_objectId"] != 0xFFFFFFFF).mean() > 0.05


def test_oracle_volume_matches_fixture():
    _, (sc_v, vol, cam_v) = MF.fixture_scenes()
    img, ns = orc.volume_image(vol, cam_v, MF.W, MF.H, inv_sampling_rate=1.0)
    np.testing.assert_array_equal(img, FIX["volume_rgba"])
    assert ns == int(FIX["volume_samples"][0])
    o = orc.OracleScene(sc_v).render(cam_v, MF.W, MF.H, volumeSamplingRate=1.0)
    for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(o[ch], FIX[f"volume_frame_{ch}"], err_msg=ch)
    # the flag's bits name the kind of tie; the fixture records which pixels
    np.testing.assert_array_equal(o["ambiguous"] != 0, FIX["volume_frame_ambiguous"] != 0)
    # the frame composites the integral over the default black background
    # after the raycast renderer's second multiply by opacity
    # (Raycast_ptx.cu:159): rgb = C.rgb * O, alpha = O + (1 - O) * 1
    exp_rgb = (img[..., :3] * img[..., 3:4]).astype(np.float32)
    np.testing.assert_allclose(o["color"][..., :3], exp_rgb, rtol=0, atol=2e-6)
    np.testing.assert_allclose(o["color"][..., 3], 1.0, atol=1e-6)
