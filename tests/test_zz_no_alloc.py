"""The allocation-free batched entry point (vrx_render_batch_static, no
counterpart in the reference) against vrx_render_batch, through the C ABI.

Two calls in a row, the second on a smaller input (fewer views, a smaller
image): every channel of every view must be bit-identical to the allocating
entry point's output for the same cameras, and the static variant must not
touch the caller's (absent) workspace or queue counter.
"""
import ctypes as C

import numpy as np
import pytest

from visrtx_amd import _native as N
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CHANNELS = ["color", "depth", "primitiveId", "objectId", "instanceId"]


def _scene():
    sc, cams = S.batched_workload(views=6, res=80, instances=24, meshes=4, tri_min=2_000,
                                  tri_max=20_000, box=4.0, seed=11)
    return sc, cams


def _static_render(B, bf):
    b = N.vrx_batch.from_buffer_copy(bf.b)
    b.d_scratch = None
    b.scratch_bytes = 0
    b.d_work_counter = None
    st = N.hip().vrx_render_batch_static(C.byref(bf.scene), C.byref(bf.renderer), C.byref(b),
                                         torch.cuda.current_stream().cuda_stream)
    N.check(st, "vrx_render_batch_static")


def _fill(bf):
    bf.depth.fill_(-2.0)
    bf.prim_id.fill_(-1)
    bf.obj_id.fill_(-1)
    bf.inst_id.fill_(-1)
    bf.color.fill_(-1)


@pytest.mark.parametrize("kind", ["raycast", "ao"])
def test_static_scratch_batch_equals_allocating_entry(require_gpu, kind):
    from visrtx_amd import batch as B
    sc, cams = _scene()
    dev = A.Device()
    s = S.build_anari(dev, sc)
    extra = {"pixelSamples": 2} if kind == "ao" else {}
    r = S.make_renderer(dev, s, kind, background=(0.1, 0.2, 0.3, 1.0), **extra)
    scene = B.scene_of(dev, s.world, wide=True)
    ren = B.renderer_of(dev, r)
    for call, (views, w, h) in enumerate([(6, 80, 60), (3, 40, 28)]):
        print(f"[no-alloc] {kind} call {call}: {views} views {w}x{h}", flush=True)
        sel = cams[:views]
        for c in sel:
            c.aspect = w / h
        ref = B.BatchFrame(scene, ren, sel, w, h)
        got = B.BatchFrame(scene, ren, sel, w, h)
        for bf in (ref, got):
            bf.b.tile_w, bf.b.tile_h, bf.b.block_threads = 16, 4, 128
            bf.renderer.node_width = 4
            _fill(bf)
        ref.render()
        _static_render(B, got)
        torch.cuda.synchronize()
        a, g = ref.channels(), got.channels()
        for ch in CHANNELS:
            np.testing.assert_array_equal(g[ch], a[ch], err_msg=f"call {call}: {ch}")
        assert (a["depth"] < 1e29).mean() > 0.05
    assert not dev.errors(), dev.errors()
    s.release()
    dev.close()


def test_static_scratch_batch_refuses_bad_arguments(require_gpu):
    from visrtx_amd import batch as B
    sc, cams = _scene()
    dev = A.Device()
    s = S.build_anari(dev, sc)
    r = S.make_renderer(dev, s, "raycast")
    bf = B.BatchFrame(B.scene_of(dev, s.world), B.renderer_of(dev, r), cams[:2], 16, 16)
    b = N.vrx_batch.from_buffer_copy(bf.b)
    b.tile_w = 7                          # a tile must hold 64 pixels
    st = N.hip().vrx_render_batch_static(C.byref(bf.scene), C.byref(bf.renderer), C.byref(b),
                                         torch.cuda.current_stream().cuda_stream)
    assert st != 0
    s.release()
    dev.close()
