"""Batched multi-view rendering (one persistent launch for every view) vs the
per-view ANARI frame and the CPU oracle.

Bar: every channel of every view bit-identical to the per-view raycast frame
of the same camera (same per-pixel program, same BVHs), and the ids exact /
depth and colour within 1e-4 relative of the oracle.  HIP-graph replays are
checked one by one against the eager result, with eager torch work and
synchronisations between the replays.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CHANNELS = ["color", "depth", "primitiveId", "objectId", "instanceId"]


def _small_batched(views=6, res=(70, 45)):
    sc, cams = S.batched_workload(views=views, res=res[0], instances=24, meshes=4,
                                  tri_min=2_000, tri_max=20_000, box=4.0, seed=11)
    for c in cams:
        c.aspect = res[0] / res[1]
    return sc, cams


def _setup(sc):
    from visrtx_amd import batch as B
    dev = A.Device()
    s = S.build_anari(dev, sc)
    r = S.make_renderer(dev, s, "raycast", background=(0.1, 0.2, 0.3, 1.0))
    return B, dev, s, r


def _per_view(dev, s, r, cam, w, h):
    c = S.make_camera(dev, s, cam)
    f = S.make_frame(dev, s, s.world, c, r, w, h)
    dev.render(f)
    dev.ready(f)
    out = {ch: dev.map_frame(f, "channel." + ch)[0].copy() for ch in CHANNELS}
    s.drop(f)
    s.drop(c)
    return out


def _assert_same(a, b, what):
    for ch in CHANNELS:
        np.testing.assert_array_equal(a[ch], b[ch], err_msg=f"{what}: {ch}")


def test_batched_equals_per_view_frames_and_oracle(require_gpu):
    w, h = 70, 45
    sc, cams = _small_batched(6, (w, h))
    B, dev, s, r = _setup(sc)
    scene = B.scene_of(dev, s.world)
    assert scene.num_entries == 24
    bf = B.BatchFrame(scene, B.renderer_of(dev, r), cams, w, h)
    print(f"[batch] {len(cams)} views {w}x{h}, 24 instances", flush=True)
    for tw, th, block in [(8, 8, 256), (64, 1, 64), (16, 4, 128), (32, 2, 256)]:
        bf.b.tile_w, bf.b.tile_h, bf.b.block_threads = tw, th, block
        bf.depth.fill_(-1.0)
        bf.render()
        torch.cuda.synchronize()
        got = bf.channels()
        for v, cam in enumerate(cams):
            ref = _per_view(dev, s, r, cam, w, h)
            _assert_same({k: got[k][v] for k in CHANNELS}, ref, f"view {v} tile {tw}x{th}")
    assert not dev.errors(), dev.errors()
    osc = orc.OracleScene(sc)
    for v in (0, 3):
        o = osc.render(cams[v], w, h, background=(0.1, 0.2, 0.3, 1.0))
        np.testing.assert_array_equal(got["primitiveId"][v], o["primitiveId"])
        np.testing.assert_array_equal(got["objectId"][v], o["objectId"])
        np.testing.assert_array_equal(got["instanceId"][v], o["instanceId"])
        np.testing.assert_allclose(got["depth"][v], o["depth"], rtol=1e-4, atol=0)
        np.testing.assert_allclose(got["color"][v], o["color"], rtol=1e-4, atol=1e-4)
    assert len(np.unique(got["instanceId"])) > 8
    bf = None
    s.release()
    dev.close()


def test_batched_graph_replays_and_autotune(require_gpu):
    w, h = 64, 64
    sc, cams = _small_batched(5, (w, h))
    B, dev, s, r = _setup(sc)
    bf = B.BatchFrame(B.scene_of(dev, s.world), B.renderer_of(dev, r), cams, w, h, srgb=True)
    bf.render()
    torch.cuda.synchronize()
    eager = bf.channels()
    best = bf.autotune()
    print(f"[batch] autotune best {best:.3f} ms: tile {bf.b.tile_w}x{bf.b.tile_h} "
          f"block {bf.b.block_threads} stack {bf.renderer.stack_lds}", flush=True)
    assert bf.b.tile_w * bf.b.tile_h == 64 and best > 0
    stream = torch.cuda.Stream()
    bf.capture(stream)
    junk = torch.randn(1 << 20, device="cuda")
    for rep in range(4):
        for t in (bf.depth, bf.prim_id, bf.obj_id, bf.inst_id, bf.color):
            t.zero_()
        junk = junk * 1.0001 + 1.0          # eager work between replays
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            bf.replay(stream)
        stream.synchronize()
        _assert_same(bf.channels(), eager, f"graph replay {rep}")
    bf.release_graph()
    s.release()
    dev.close()


def test_batched_rejects_bad_tiles_and_small_scratch(require_gpu):
    import ctypes as C
    from visrtx_amd import _native as N
    sc, cams = _small_batched(2, (32, 32))
    B, dev, s, r = _setup(sc)
    bf = B.BatchFrame(B.scene_of(dev, s.world), B.renderer_of(dev, r), cams, 32, 32)
    L = N.hip()
    bf.b.tile_w, bf.b.tile_h = 8, 4
    assert L.vrx_render_batch(C.byref(bf.scene), C.byref(bf.renderer), C.byref(bf.b), None) \
        == -2
    bf.b.tile_w, bf.b.tile_h = 8, 8
    bf.b.scratch_bytes = 16
    assert L.vrx_render_batch(C.byref(bf.scene), C.byref(bf.renderer), C.byref(bf.b), None) \
        == -2
    assert b"scratch" in L.vrx_last_error()
    # an unknown renderer kind is refused
    rr = N.vrx_renderer.from_buffer_copy(bf.renderer)
    rr.kind = 7
    assert L.vrx_render_batch(C.byref(bf.scene), C.byref(rr), C.byref(bf.b), None) == -2
    assert b"renderer kind" in L.vrx_last_error()
    s.release()
    dev.close()


def test_batched_wide_nodes_equal_per_view(require_gpu):
    """Node width 4 in the persistent launch: same channels as the BVH2
    per-view frames; the autotuner may pick either width."""
    w, h = 64, 48
    sc, cams = _small_batched(4, (w, h))
    B, dev, s, r = _setup(sc)
    scene = B.scene_of(dev, s.world, wide=True)
    assert scene.wide == 1
    bf = B.BatchFrame(scene, B.renderer_of(dev, r), cams, w, h)
    bf.renderer.node_width = 4
    bf.render()
    torch.cuda.synchronize()
    got = bf.channels()
    for v, cam in enumerate(cams):
        ref = _per_view(dev, s, r, cam, w, h)
        _assert_same({k: got[k][v] for k in CHANNELS}, ref, f"view {v} width 4")
    bf.autotune()
    assert bf.renderer.node_width in (2, 4)
    bf.render()
    torch.cuda.synchronize()
    _assert_same(bf.channels(), got, "autotuned")
    s.release()
    dev.close()


@pytest.mark.parametrize("kind", ["ao", "directLight"])
def test_batched_ao_and_direct_light_equal_per_view_frames(require_gpu, kind):
    """The AO and directLight programs in the persistent batched launch: each
    view equals the same view rendered alone (same per-pixel PCG32 sequence)."""
    w, h = 48, 32
    sc, cams = _small_batched(4, (w, h))
    sc.lights = [S.LightDesc("directional", direction=(0.3, -1.0, 0.2), irradiance=1.5),
                 S.LightDesc("point", position=(0.0, 3.0, 0.0), intensity=1.0)]
    from visrtx_amd import batch as B
    dev = A.Device()
    s = S.build_anari(dev, sc)
    r = S.make_renderer(dev, s, kind, pixelSamples=2, ambientSamples=2,
                        ambientOcclusionDistance=0.5, ambientRadiance=0.3)
    bf = B.BatchFrame(B.scene_of(dev, s.world), B.renderer_of(dev, r), cams, w, h)
    print(f"[batch] {kind} {len(cams)} views {w}x{h}", flush=True)
    bf.render()
    torch.cuda.synchronize()
    got = bf.channels()
    for v, cam in enumerate(cams):
        ref = _per_view(dev, s, r, cam, w, h)
        _assert_same({k: got[k][v] for k in CHANNELS}, ref, f"{kind} view {v}")
    assert not dev.errors(), dev.errors()
    assert got["color"][..., :3].max() > 0.05
    bf = None
    s.release()
    dev.close()
