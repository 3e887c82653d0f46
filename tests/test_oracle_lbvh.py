"""Pins the CPU oracle's LBVH build (oracle/oracle.c) before it is trusted.

- Morton codes: hand-computed known answers on a unit centroid box, and the
  bit-interleave closed form.
- Sort: equality with numpy's stable lexsort of (code, index).
- Karras hierarchy: equality with the binary radix tree defined top-down
  (each node splits its key range where the highest differing bit of its
  end keys flips), plus coverage invariants.
- Refit: every node box equals the min/max over the leaves of its subtree.
"""
import numpy as np
import pytest

from oracle import oracle as orc

LEAF = orc.LEAF


def spread_bits(v: int) -> int:
    r = 0
    for i in range(10):
        if v >> i & 1:
            r |= 1 << (3 * i)
    return r


def test_expand_bits_closed_form():
    rng = np.random.default_rng(0)
    for v in list(rng.integers(0, 1024, 200)) + [0, 1, 512, 1023]:
        assert orc.expand_bits10(int(v)) == spread_bits(int(v))


def test_morton_known_answers():
    # corner spheres fix the centroid box to [0,1]^3 exactly (scale = 1024)
    centers = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.25, 0.75], [0.999, 0.0, 0.5],
                        [1.0 / 1024, 2.0 / 1024, 3.0 / 1024]], np.float32)
    boxes = orc.bounds(orc.ORC_SPHERE, centers, radius=np.full(5, 0.01, np.float32))
    codes, cbox = orc.morton(boxes)
    np.testing.assert_array_equal(cbox, [0, 0, 0, 1, 1, 1])

    def code(q):
        return spread_bits(q[0]) * 4 + spread_bits(q[1]) * 2 + spread_bits(q[2])

    expect = [code((0, 0, 0)), code((1023, 1023, 1023)), code((512, 256, 768)),
              code((1022, 0, 512)), code((1, 2, 3))]
    assert list(codes) == expect


def test_morton_degenerate_axis_and_nan():
    # all centroids share z -> scale 0 on z; a NaN centroid is ignored in the box
    c = np.array([[0, 0, 5], [1, 1, 5], [0.5, 0.5, 5]], np.float32)
    boxes = orc.bounds(orc.ORC_SPHERE, c, radius=np.zeros(3, np.float32))
    boxes = np.vstack([boxes, np.full((1, 6), np.nan, np.float32)])
    codes, cbox = orc.morton(boxes)
    np.testing.assert_array_equal(cbox, [0, 0, 5, 1, 1, 5])
    assert codes[3] == 0 and codes[0] == 0
    assert codes[1] == spread_bits(1023) * 4 + spread_bits(1023) * 2


@pytest.mark.parametrize("n", [1, 2, 5, 257, 10000])
def test_sort_is_stable_lexsort(n):
    rng = np.random.default_rng(n)
    codes = rng.integers(0, 64, n).astype(np.uint32) * np.uint32(0x01010101)
    sc, order = orc.sort(codes)
    ref = np.lexsort((np.arange(n), codes))
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(sc, codes[ref])


def _aug(sc, i):
    return (int(sc[i]) << 32) | i


def topdown_radix_tree(sc):
    """Binary radix tree over index-augmented keys, built top-down.
    Returns {(lo, hi): split} for every internal range."""
    n = len(sc)
    out = {}
    stack = [(0, n - 1)]
    while stack:
        lo, hi = stack.pop()
        if lo == hi:
            continue
        a, b = _aug(sc, lo), _aug(sc, hi)
        hb = (a ^ b).bit_length() - 1          # highest differing bit
        # first index whose key has bit hb set (keys are sorted)
        split = lo
        for k in range(lo, hi + 1):
            if _aug(sc, k) >> hb & 1:
                split = k - 1
                break
        out[(lo, hi)] = split
        stack.append((lo, split))
        stack.append((split + 1, hi))
    return out


def karras_ranges(child, n):
    """Walk the Karras output from the root: {(lo, hi): split} and visit counts.
    In Karras' numbering the left child (leaf or internal) has index gamma =
    split and the right child index gamma + 1."""
    out = {}
    seen_int = np.zeros(max(n - 1, 1), np.int64)
    seen_leaf = np.zeros(n, np.int64)
    stack = [(0, 0, n - 1)]
    while stack:
        node, lo, hi = stack.pop()
        seen_int[node] += 1
        l, r = (int(x) for x in child[node])
        split = l & ~LEAF
        assert (r & ~LEAF) == split + 1
        assert lo <= split < hi
        out[(lo, hi)] = split
        if l & LEAF:
            assert split == lo
            seen_leaf[split] += 1
        else:
            stack.append((split, lo, split))
        if r & LEAF:
            assert split + 1 == hi
            seen_leaf[split + 1] += 1
        else:
            stack.append((split + 1, split + 1, hi))
    return out, seen_int, seen_leaf


@pytest.mark.parametrize("n,dup", [(2, False), (3, False), (9, False), (64, True), (1000, False),
                                   (1000, True), (4097, True)])
def test_karras_equals_topdown_radix_tree(n, dup):
    rng = np.random.default_rng(n + dup)
    hi = 8 if dup else 1 << 30
    codes = np.sort(rng.integers(0, hi, n).astype(np.uint32))
    child, parent = orc.karras(codes)
    kr, seen_int, seen_leaf = karras_ranges(child, n)
    td = topdown_radix_tree(codes)
    assert kr == td
    assert (seen_int == 1).all(), "every internal node reached exactly once"
    assert (seen_leaf == 1).all(), "every leaf reached exactly once"
    # parent links agree with children
    for i, (l, r) in enumerate(child):
        for c in (int(l), int(r)):
            pi = (n - 1) + (c & ~LEAF) if c & LEAF else c
            assert parent[pi] == i


@pytest.mark.parametrize("type_", [orc.ORC_SPHERE, orc.ORC_TRIANGLE, orc.ORC_CYLINDER])
def test_refit_is_subtree_union(type_):
    rng = np.random.default_rng(7 + type_)
    n = 600
    if type_ == orc.ORC_SPHERE:
        pos = rng.random((n, 3), dtype=np.float32)
        rad = rng.uniform(0.001, 0.02, n).astype(np.float32)
        boxes = orc.bounds(type_, pos, radius=rad)
    elif type_ == orc.ORC_TRIANGLE:
        pos = rng.random((3 * n, 3), dtype=np.float32)
        boxes = orc.bounds(type_, pos)
    else:
        pos = rng.random((2 * n, 3), dtype=np.float32)
        rad = rng.uniform(-0.02, 0.02, n).astype(np.float32)   # |r| is used
        boxes = orc.bounds(type_, pos, radius=rad)
    codes, _ = orc.morton(boxes)
    sc, order = orc.sort(codes)
    child, _ = orc.karras(sc)
    nb, root = orc.refit(boxes, order, child)
    kr, _, _ = karras_ranges(child, n)
    # build node -> range map
    ranges = {}
    stack = [(0, 0, n - 1)]
    while stack:
        node, lo, hi_ = stack.pop()
        ranges[node] = (lo, hi_)
        split = kr[(lo, hi_)]
        l, r = (int(x) for x in child[node])
        if not (l & LEAF):
            stack.append((int(l), lo, split))
        if not (r & LEAF):
            stack.append((int(r), split + 1, hi_))
    sb = boxes[order]
    for node, (lo, hi_) in ranges.items():
        split = kr[(lo, hi_)]
        left = sb[lo:split + 1]
        right = sb[split + 1:hi_ + 1]
        np.testing.assert_array_equal(nb[node, 0:3], left[:, 0:3].min(0))
        np.testing.assert_array_equal(nb[node, 3:6], left[:, 3:6].max(0))
        np.testing.assert_array_equal(nb[node, 6:9], right[:, 0:3].min(0))
        np.testing.assert_array_equal(nb[node, 9:12], right[:, 3:6].max(0))
    np.testing.assert_array_equal(root[:3], boxes[:, :3].min(0))
    np.testing.assert_array_equal(root[3:], boxes[:, 3:].max(0))


def test_cylinder_bounds_use_abs_radius():
    pos = np.array([[0, 0, 0], [1, 2, 3]], np.float32)
    b = orc.bounds(orc.ORC_CYLINDER, pos, radius=np.array([-0.5], np.float32))
    np.testing.assert_array_equal(b[0], [-0.5, -0.5, -0.5, 1.5, 2.5, 3.5])
