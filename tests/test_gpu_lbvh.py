"""LBVH build on the GPU vs the CPU oracle: bit-exact Morton codes, sort
order, Karras topology, refit boxes and root header (through the C ABI
vrx_bvh_build_debug / vrx_bvh_build)."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from tests.devmem import DeviceBuffer
from visrtx_amd import _native as N
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

LEAF = 0x80000000


def gpu_build(type_, pos, index=None, radius=None, global_radius=0.0):
    L = N.hip()
    n = orc._num_prims(type_, pos, index)
    sz = N.vrx_bvh_sizes()
    N.check(L.vrx_bvh_query_sizes(type_, C.c_uint32(n), C.byref(sz)))
    dpos = DeviceBuffer.from_array(np.ascontiguousarray(pos, np.float32))
    didx = DeviceBuffer.from_array(np.ascontiguousarray(index, np.uint32)) if index is not None else None
    drad = DeviceBuffer.from_array(np.ascontiguousarray(radius, np.float32)) if radius is not None else None
    nodes, leaves, ids = DeviceBuffer(sz.nodes), DeviceBuffer(sz.leaves), DeviceBuffer(sz.leaf_ids)
    header, ws = DeviceBuffer(sz.header), DeviceBuffer(sz.workspace)
    codes, cbox = DeviceBuffer(4 * max(n, 1)), DeviceBuffer(64)
    p = N.vrx_prims(type_, n, dpos.addr, didx.addr if didx else None, drad.addr if drad else None,
                    global_radius)
    b = N.vrx_bvh(nodes.addr, leaves.addr, ids.addr, header.addr)
    print(f"[lbvh] type={type_} n={n}", flush=True)
    N.check(L.vrx_bvh_build_debug(C.byref(p), C.byref(b), ws.addr, C.c_size_t(sz.workspace),
                                  codes.addr, cbox.addr, None), "vrx_bvh_build_debug")
    N.check(L.vrx_device_sync())
    out = {}
    out["codes"] = codes.to_array(np.uint32, n)
    out["cbox"] = cbox.to_array(np.float32, 6)
    out["ids"] = ids.to_array(np.uint32, n)
    raw = nodes.to_array(np.uint32, 16 * max(n - 1, 0))
    raw = raw.reshape(-1, 16)
    out["child"] = raw[:, 12:14].copy()
    out["heights"] = raw[:, 14:16].copy()
    out["boxes"] = raw[:, 0:12].copy().view(np.float32)
    h = header.to_array(np.uint32, 16)
    out["root_box"] = h[0:6].view(np.float32)
    out["height"] = int(h[6])
    out["root"] = int(h[7])
    out["leaves"] = leaves.to_array(np.float32, (sz.leaves // 4))
    return out


def oracle_build(type_, pos, index=None, radius=None, global_radius=0.0):
    boxes = orc.bounds(type_, pos, index, radius, global_radius)
    codes, cbox = orc.morton(boxes)
    sc, order = orc.sort(codes)
    child, _ = orc.karras(sc)
    nb, root = orc.refit(boxes, order, child)
    return dict(codes=sc, cbox=cbox, ids=order, child=child, boxes=nb, root_box=root, raw=boxes)


def tree_height(child, n):
    """Max leaf depth + 1, by levels (root = level 1)."""
    if n == 1:
        return 1
    level = np.array([0], np.int64)
    h = 1
    while level.size:
        ch = child[level].ravel().astype(np.int64)
        level = ch[(ch & LEAF) == 0]
        h += 1
    return h


def check_equal(g, o, n):
    np.testing.assert_array_equal(g["cbox"], o["cbox"])
    np.testing.assert_array_equal(g["codes"], o["codes"])
    np.testing.assert_array_equal(g["ids"], o["ids"])
    if n >= 2:
        np.testing.assert_array_equal(g["child"], o["child"])
        np.testing.assert_array_equal(g["boxes"], o["boxes"])
    np.testing.assert_array_equal(g["root_box"], o["root_box"])
    assert g["height"] == tree_height(o["child"], n)
    assert g["root"] == (LEAF if n == 1 else 0)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 4096, 4097, 100_000, 1_000_000])
def test_spheres_bit_exact(require_gpu, n):
    c, r = S.gen_spheres(n, seed=11 + n)
    g = gpu_build(N.VRX_SPHERE, c, radius=r)
    o = oracle_build(orc.ORC_SPHERE, c, radius=r)
    check_equal(g, o, n)
    # leaf records are the spheres in sorted order
    lv = g["leaves"][: 4 * n].reshape(n, 4)
    np.testing.assert_array_equal(lv[:, :3], c[o["ids"]])
    np.testing.assert_array_equal(lv[:, 3], r[o["ids"]])


def test_duplicate_centroids_bit_exact(require_gpu):
    # 5000 spheres on only 7 distinct centers: index-augmented Karras keys
    rng = np.random.default_rng(5)
    base = rng.random((7, 3), dtype=np.float32)
    c = base[rng.integers(0, 7, 5000)]
    r = np.full(5000, 0.01, np.float32)
    check_equal(gpu_build(N.VRX_SPHERE, c, radius=r), oracle_build(orc.ORC_SPHERE, c, radius=r), 5000)


def test_indexed_spheres_global_radius(require_gpu):
    c, _ = S.gen_spheres(3000, seed=3)
    idx = np.random.default_rng(2).integers(0, 3000, 2000).astype(np.uint32)
    check_equal(gpu_build(N.VRX_SPHERE, c, index=idx, global_radius=0.02),
                oracle_build(orc.ORC_SPHERE, c, index=idx, global_radius=0.02), 2000)


def test_heightfield_triangles_bit_exact(require_gpu):
    v, idx = S.gen_heightfield(257, seed=1)
    n = len(idx)
    g = gpu_build(N.VRX_TRIANGLE, v, index=idx)
    o = oracle_build(orc.ORC_TRIANGLE, v, index=idx)
    check_equal(g, o, n)
    lv = g["leaves"][: 12 * n].reshape(n, 3, 4)
    tri = v[idx[o["ids"]]]
    np.testing.assert_array_equal(lv[:, :, :3], tri)


def test_triangle_soup_and_cylinders(require_gpu):
    rng = np.random.default_rng(9)
    soup = rng.random((3 * 777, 3), dtype=np.float32)
    check_equal(gpu_build(N.VRX_TRIANGLE, soup), oracle_build(orc.ORC_TRIANGLE, soup), 777)
    cv = rng.random((2 * 555, 3), dtype=np.float32)
    cr = rng.uniform(-0.05, 0.05, 555).astype(np.float32)
    check_equal(gpu_build(N.VRX_CYLINDER, cv, radius=cr), oracle_build(orc.ORC_CYLINDER, cv, radius=cr), 555)


@pytest.mark.slow
def test_10m_spheres_bit_exact(require_gpu):
    """The 10M-sphere benchmark scene, full size."""
    c, r = S.gen_spheres(10_000_000, seed=2)
    g = gpu_build(N.VRX_SPHERE, c, radius=r)
    o = oracle_build(orc.ORC_SPHERE, c, radius=r)
    check_equal(g, o, 10_000_000)


def test_wide_nodes_match_bvh2_collapse(require_gpu):
    """vrx_bvh_widen: record i holds node i's grandchildren (a leaf child keeps
    its slot), boxes copied bit for bit from the BVH2 nodes."""
    from tests.devmem import DeviceBuffer as DB
    c, r = S.gen_spheres(5000, seed=12)
    L = N.hip()
    n = len(c)
    sz = N.vrx_bvh_sizes()
    N.check(L.vrx_bvh_query_sizes(N.VRX_SPHERE, C.c_uint32(n), C.byref(sz)))
    dpos, drad = DB.from_array(c), DB.from_array(r)
    nodes, leaves, ids = DB(sz.nodes), DB(sz.leaves), DB(sz.leaf_ids)
    header, ws = DB(sz.header), DB(sz.workspace)
    wide = DB(L.vrx_bvh_wide_bytes(n))
    p = N.vrx_prims(N.VRX_SPHERE, n, dpos.addr, None, drad.addr, 0.0)
    b = N.vrx_bvh(nodes.addr, leaves.addr, ids.addr, header.addr, wide.addr)
    N.check(L.vrx_bvh_build(C.byref(p), C.byref(b), ws.addr, C.c_size_t(sz.workspace), None))
    N.check(L.vrx_bvh_widen(C.byref(b), n, None), "vrx_bvh_widen")
    N.check(L.vrx_device_sync())
    raw = nodes.to_array(np.uint32, 16 * (n - 1)).reshape(-1, 16)
    w = wide.to_array(np.uint32, 32 * (n - 1)).reshape(-1, 32)
    boxes = raw[:, 0:12]
    child = raw[:, 12:14]

    def kids(i):
        out = []
        for j in range(2):
            bx = boxes[i, 6 * j:6 * j + 6]
            lk = int(child[i, j])
            if lk & LEAF:
                out.append((bx, lk))
            else:
                for g in range(2):
                    out.append((boxes[lk, 6 * g:6 * g + 6], int(child[lk, g])))
        return out

    for i in range(n - 1):
        k = kids(i)
        assert w[i, 28] == len(k)
        for s_, (bx, lk) in enumerate(k):
            got = [w[i, 4 * comp + s_] for comp in range(6)]
            np.testing.assert_array_equal(np.array(got, np.uint32), bx, err_msg=f"node {i}")
            assert w[i, 24 + s_] == lk
