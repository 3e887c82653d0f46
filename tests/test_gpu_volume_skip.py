"""Macro-cell majorant grid (UniformGrid.cu:44-258: 16^3-voxel cells) and the
empty-space skipping it drives in the volume march.

The grid is checked bit for bit against a numpy restatement of its rule
(each cell's value range over its voxels plus one on every side, then the
largest LUT alpha over the entries that range classifies to).  Skipping is
exact: a sample whose LUT alpha is 0 has step transmittance exactly 1 and
adds nothing, so frames with skipping equal frames without it bit for bit,
and both match the FP64 oracle as the other volume tests do; only the
volumeSamples counter drops."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from tests.devmem import DeviceBuffer
from tests.test_gpu_render import render_gpu
from tests.test_oracle_majorants import grid_reference as _grid_reference
from visrtx_amd import _native as N
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("vr", [(0.0, 1.0), (0.7, 0.3)])
@pytest.mark.parametrize("shape", [(40, 33, 50), (16, 16, 16), (1, 20, 17)])
def test_majorant_grid_matches_restatement(require_gpu, shape, vr):
    """A reversed value range classifies every value to the last texel
    (position()'s glm::clamp, gpu_math.h:185-189), so every cell with a
    finite voxel gets that texel's alpha."""
    rng = np.random.default_rng(sum(shape))
    data = rng.random(shape, dtype=np.float32)
    data[data < 0.6] = 0.0
    data[:18, :18, :18] = 0.0                    # cell (0, 0, 0) and its halo: empty
    data[tuple(rng.integers(0, s, 5) for s in shape)] = np.nan
    lut = orc.tf_discretize(S.COLORMAP5, np.array([0.0, 0.0, 0.3, 1.0], np.float32), (0.0, 1.0))
    want = _grid_reference(data, lut, vr)
    L = N.hip()
    db, lb = DeviceBuffer.from_array(data), DeviceBuffer.from_array(lut)
    v = N.vrx_volume()
    v.d_data, v.d_lut = db.addr, lb.addr
    v.dims[:] = [shape[2], shape[1], shape[0]]
    v.spacing[:] = [1.0, 1.0, 1.0]
    v.value_range[:] = list(vr)
    v.unit_distance = 1.0
    cells = L.vrx_volume_grid_cells(C.byref(v))
    assert cells == want.size
    out = DeviceBuffer(4 * cells)
    print(f"[grid] shape {shape} cells {cells}", flush=True)
    N.check(L.vrx_volume_majorant_grid(C.byref(v), out.ptr, None), "grid")
    N.check(L.vrx_device_sync(), "sync")
    got = out.to_array(np.float32, cells).reshape(want.shape)
    np.testing.assert_array_equal(got, want)
    if vr[0] < vr[1]:
        assert want.reshape(-1)[0] == 0 and (want.size == 1 or (want > 0).any())
    else:
        assert (want == lut.reshape(-1, 4)[254:, 3].max()).all()
    for b in (db, lb, out):
        b.free()


def _render_both(sc, cam, w, h, kind, **rp):
    on = render_gpu(sc, cam, w, h, kind=kind, emptySpaceSkipping=True, **rp)
    off = render_gpu(sc, cam, w, h, kind=kind, emptySpaceSkipping=False, **rp)
    for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(on[ch], off[ch], err_msg=ch)
    return on, off


def _vol_samples(sc, cam, w, h, kind, skip, **rp):
    from visrtx_amd import anari as A
    dev = A.Device()
    s = S.build_anari(dev, sc)
    c = S.make_camera(dev, s, cam)
    r = S.make_renderer(dev, s, kind, emptySpaceSkipping=skip, **rp)
    f = S.make_frame(dev, s, s.world, c, r, w, h)
    dev.render(f)
    dev.ready(f)
    n = dev.get_property(f, "volumeSamples", A.UINT64)
    s.release()
    dev.close()
    return n


def test_skipping_is_exact_volume_only(require_gpu):
    """Volume-only raycast (the LDS-LUT kernel): same frame with and without
    skipping, equal to the FP64 oracle, with far fewer samples taken."""
    sc, cam = S.sparse_volume_workload(128, width=200, height=120)
    on, _ = _render_both(sc, cam, 200, 120, "raycast", volumeSamplingRate=0.5)
    o = orc.OracleScene(sc).render(cam, 200, 120, volumeSamplingRate=0.5)
    ok = o["ambiguous"] == 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(on["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    np.testing.assert_array_equal(on["objectId"], o["objectId"])
    n_on = _vol_samples(sc, cam, 200, 120, "raycast", True, volumeSamplingRate=0.5)
    n_off = _vol_samples(sc, cam, 200, 120, "raycast", False, volumeSamplingRate=0.5)
    print(f"volume samples: {n_on} with skipping, {n_off} without")
    assert n_on < 0.5 * n_off


@pytest.mark.parametrize("vr,opacity", [((0.25, 0.6), None), ((0.6, 0.25), (0.8, 0.4, 0.0, 0.0))])
def test_skipping_is_exact_for_value_ranges(require_gpu, vr, opacity):
    """A narrow value range (values below it clamp to the first texel, alpha
    0, so cells stay skippable) and a reversed one.  A reversed range
    discretizes the transfer function to its first colour and opacity with a
    NaN entry 0 (TransferFunction1D.cpp:101-150), and every sample reads the
    last texel, so no cell is empty and the world marches the field with
    the dense four-lane kernel instead.  Frames equal with and without
    skipping and match the FP64 oracle."""
    sc, cam = S.sparse_volume_workload(96, width=160, height=100)
    sc.volumes[0].value_range = vr
    if opacity is not None:
        sc.volumes[0].opacity = np.array(opacity, np.float32)
    on, _ = _render_both(sc, cam, 160, 100, "raycast", volumeSamplingRate=0.5)
    o = orc.OracleScene(sc).render(cam, 160, 100, volumeSamplingRate=0.5)
    ok = o["ambiguous"] == 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(on["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    np.testing.assert_array_equal(on["objectId"], o["objectId"])
    np.testing.assert_allclose(on["depth"], o["depth"], rtol=1e-5)
    hit = on["depth"] < 1e29
    assert hit.mean() > 0.05 and (on["color"][hit][:, :3].sum(axis=1) > 0).mean() > 0.05


def test_skipping_is_exact_with_surfaces_and_lights(require_gpu):
    """The generic kernel (march_all before surfaces) and directLight's
    volume shadow rays skip too, with identical frames."""
    sc, cam = S.sparse_volume_workload(96, width=160, height=100)
    c, r = S.gen_spheres(400, seed=4, rmin=0.01, rmax=0.03)
    sc.geoms = [S.GeomDesc("sphere", c, radius=r, surface_id=3)]
    sc.world_geoms = [0]
    sc.lights = [S.LightDesc("directional", direction=(0.2, -1.0, 0.3), irradiance=2.0)]
    _render_both(sc, cam, 160, 100, "ao", pixelSamples=1, ambientSamples=1,
                 volumeSamplingRate=0.5)
    on, _ = _render_both(sc, cam, 160, 100, "directLight", pixelSamples=1, ambientSamples=1,
                         ambientRadiance=0.2, volumeSamplingRate=0.5)
    o = orc.OracleScene(sc).render(cam, 160, 100, kind="directLight", pixelSamples=1,
                                   ambientSamples=1, ambientRadiance=0.2, volumeSamplingRate=0.5)
    ok = o["ambiguous"] == 0
    assert ok.mean() > 0.98
    np.testing.assert_allclose(on["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    np.testing.assert_array_equal(on["primitiveId"], o["primitiveId"])


def test_woodcock_sampling_matches_oracle(require_gpu):
    """vrx_volume_sample_distance against the oracle's restatement on the same
    rays and PCG32 streams: the collision flag and distance agree on all but
    the rare rays whose fp32 / FP64 field samples fall on opposite sides of
    an acceptance threshold."""
    n = 80
    data = S.gen_noise_volume(n, seed=3)
    data[:, :, : n // 3] = 0.0                                 # an empty third
    sp = 1.0 / (n - 1)
    vol = S.VolumeDesc(data, spacing=(sp, sp, sp), color=S.COLORMAP5,
                       opacity=np.array([0.0, 0.0, 0.05, 0.3], np.float32))
    lut = orc.tf_discretize(vol.color, vol.opacity, vol.value_range)
    rng = np.random.default_rng(4)
    m = 20000
    org = (rng.random((m, 3)) * 3.0 - 1.0).astype(np.float32)
    tgt = rng.random((m, 3)).astype(np.float32)
    dirs = tgt - org
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    tmm = np.tile(np.array([0.0, 1e30], np.float32), (m, 1))
    tmm[::7, 1] = 1.5                                          # some bounded rays
    L = N.hip()
    db, lb = DeviceBuffer.from_array(data), DeviceBuffer.from_array(lut)
    v = N.vrx_volume()
    v.d_data, v.d_lut = db.addr, lb.addr
    v.dims[:] = [n, n, n]
    v.spacing[:] = [sp, sp, sp]
    v.value_range[:] = [0.0, 1.0]
    v.unit_distance = 1.0
    cells = L.vrx_volume_grid_cells(C.byref(v))
    gb = DeviceBuffer(4 * cells)
    N.check(L.vrx_volume_majorant_grid(C.byref(v), gb.ptr, None), "grid")
    v.d_majorant = gb.addr
    bufs = [DeviceBuffer.from_array(a) for a in (org, dirs, tmm)]
    out_t, out_a, out_tr = DeviceBuffer(4 * m), DeviceBuffer(12 * m), DeviceBuffer(4 * m)
    print(f"[woodcock] {m} rays through {n}^3", flush=True)
    N.check(L.vrx_volume_sample_distance(C.byref(v), m, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                         C.c_uint64(9), out_t.ptr, out_a.ptr, out_tr.ptr, None),
            "sample_distance")
    N.check(L.vrx_device_sync(), "sync")
    g_t, g_tr = out_t.to_array(np.float32, m), out_tr.to_array(np.float32, m)
    g_grid = gb.to_array(np.float32, cells)
    np.testing.assert_array_equal(g_grid, orc.majorant_grid(vol).reshape(-1))
    o_t, o_a, o_tr = orc.sample_distance(vol, org, dirs, tmm, seed=9)
    same = (g_tr == o_tr) & np.isclose(g_t, o_t, rtol=1e-4, atol=1e-6)
    print(f"agree {same.mean():.5f}, collisions {(g_tr == 0).mean():.3f}")
    assert same.mean() > 0.995
    assert 0.05 < (g_tr == 0).mean() < 0.95
    for b in [db, lb, gb, out_t, out_a, out_tr] + bufs:
        b.free()


def test_skipping_march_unroll_depths_give_one_frame(require_gpu):
    """The grid kernel at volumeUnroll 1, 2, 4 and 8, with and without
    skipping: unrolling only groups the gathers and a skip lands on the same
    sample whatever the group, so every frame is bit-identical (and the
    sample counter drops only with skipping)."""
    sc, cam = S.sparse_volume_workload(96, width=160, height=100)
    frames = []
    for skip in (True, False):
        for u in (1, 2, 4, 8):
            print(f"[grid] skip {skip} volumeUnroll {u}", flush=True)
            frames.append(render_gpu(sc, cam, 160, 100, emptySpaceSkipping=skip,
                                     volumeSamplingRate=0.5, volumeUnroll=u))
    for f in frames[1:]:
        assert not f["errors"], f["errors"]
        for ch in ("color", "depth", "primitiveId", "objectId", "instanceId"):
            np.testing.assert_array_equal(f[ch], frames[0][ch], err_msg=ch)
