"""numpy restatement of the reference BVH builder.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg, never by the product package.

Follows rendering/taichi/bvh_builder.py:
  * padded triangle boxes and centroids        :363-367
  * material palette de-duplication            :370-410
  * Morton order (10 bits/axis, stable argsort) :416-443
  * count-median recursive split, preorder ids  :445-486
  * median split on the longest axis            :488-558
  * leaf-order gather of triangle attributes    :527-544
and core/bvh.py:2-48 (object median split with a stable sort, fp64 boxes)
flattened in preorder.

Output layout is fields.py:50-56: aabb_min/aabb_max (n,3) f32, left/right,
tri_start/tri_count (n,) i32 with -1/0 conventions of the reference.
"""
from __future__ import annotations

import numpy as np

MAX_LEAF_SIZE = 8     # bvh_builder.py:282
MAX_DEPTH = 32        # bvh_builder.py:283
MORTON_BITS = 10      # bvh_builder.py:285


def padded_bounds(v0, v1, v2):
    """Triangle boxes grown by 1e-4 in fp32 and centroids ((v0+v1)+v2)/3."""
    v0 = np.asarray(v0, np.float32).reshape(-1, 3)
    v1 = np.asarray(v1, np.float32).reshape(-1, 3)
    v2 = np.asarray(v2, np.float32).reshape(-1, 3)
    stacked = np.stack([v0, v1, v2], axis=1)
    eps = np.float32(1e-4)
    lo = stacked.min(axis=1) - eps
    hi = stacked.max(axis=1) + eps
    cen = (v0 + v1 + v2) / np.float32(3.0)
    return lo, hi, cen


def spread_bits(v):
    """10-bit value -> every third bit of a 30-bit word (bvh_builder.py:437-443)."""
    v = np.asarray(v, dtype=np.uint32) & np.uint32(0x3FF)
    v = (v | (v << np.uint32(16))) & np.uint32(0x030000FF)
    v = (v | (v << np.uint32(8))) & np.uint32(0x0300F00F)
    v = (v | (v << np.uint32(4))) & np.uint32(0x030C30C3)
    v = (v | (v << np.uint32(2))) & np.uint32(0x09249249)
    return v


def morton_codes(cen):
    """30-bit Morton codes of normalised centroids (bvh_builder.py:416-434)."""
    cen = np.asarray(cen, np.float32)
    lo = cen.min(axis=0)
    ext = cen.max(axis=0) - lo
    ext[ext <= 1e-12] = 1.0
    scale = np.float32((1 << MORTON_BITS) - 1)
    unit = np.clip((cen - lo) / ext, 0.0, 1.0)
    q = (unit * scale).astype(np.uint32)
    return (spread_bits(q[:, 0]) << np.uint32(2)) | (spread_bits(q[:, 1]) << np.uint32(1)) | spread_bits(q[:, 2])


def morton_order(cen):
    return np.argsort(morton_codes(cen), kind="stable").astype(np.int32)


class Tree:
    """Flattened BVH in the reference's field layout plus the leaf order."""

    def __init__(self, n_nodes):
        self.aabb_min = np.zeros((n_nodes, 3), np.float32)
        self.aabb_max = np.zeros((n_nodes, 3), np.float32)
        self.left = np.full(n_nodes, -1, np.int32)
        self.right = np.full(n_nodes, -1, np.int32)
        self.tri_start = np.full(n_nodes, -1, np.int32)
        self.tri_count = np.zeros(n_nodes, np.int32)
        self.order = np.zeros(0, np.int32)
        self.depth = 0

    @property
    def n_nodes(self):
        return len(self.left)


def _finish(recs, left, right, order, lo, hi):
    """Leaf boxes from their triangles, interior boxes from the children."""
    t = Tree(len(recs))
    t.left, t.right = left, right
    t.order = order.astype(np.int32)
    depth = np.asarray([r[2] for r in recs], np.int32)
    t.depth = int(depth.max()) if len(depth) else 0
    is_leaf = left < 0
    starts = np.asarray([r[0] for r in recs], np.int64)
    counts = np.asarray([r[1] for r in recs], np.int64)
    t.tri_start[is_leaf] = starts[is_leaf]
    t.tri_count[is_leaf] = counts[is_leaf]
    olo, ohi = lo[order], hi[order]
    leaf_ids = np.nonzero(is_leaf)[0]
    # leaves are emitted left to right, so their ranges tile [0, n) in order
    ls = starts[leaf_ids]
    t.aabb_min[leaf_ids] = np.minimum.reduceat(olo, ls, axis=0)
    t.aabb_max[leaf_ids] = np.maximum.reduceat(ohi, ls, axis=0)
    for d in range(t.depth, -1, -1):
        ids = np.nonzero((~is_leaf) & (depth == d))[0]
        if len(ids):
            t.aabb_min[ids] = np.minimum(t.aabb_min[left[ids]], t.aabb_min[right[ids]])
            t.aabb_max[ids] = np.maximum(t.aabb_max[right[ids]], t.aabb_max[left[ids]])
    return t


def build_morton(v0, v1, v2, max_leaf=MAX_LEAF_SIZE):
    """GPUBVHBuilder(method="morton"): bvh_builder.py:412-486."""
    lo, hi, cen = padded_bounds(v0, v1, v2)
    n = len(cen)
    if n == 0:
        return Tree(0)
    order = morton_order(cen)

    def split(start, count, depth, payload):
        mid = start + count // 2  # (start + end) // 2 for integer start
        return (start, mid - start, None), (mid, start + count - mid, None)

    recs, left, right = _preorder_payload(n, lambda c, d: c <= max_leaf, split, None)
    return _finish(recs, left, right, order, lo, hi)


def _longest_axis(lo, hi):
    """bvh_builder.py:552-558 and core/bvh.py:41-48: ties prefer x, then y."""
    ex, ey, ez = (np.asarray(hi, np.float64) - np.asarray(lo, np.float64)).tolist()
    if ex >= ey and ex >= ez:
        return 0
    if ey >= ez:
        return 1
    return 2


def build_median(v0, v1, v2, max_leaf=MAX_LEAF_SIZE, max_depth=MAX_DEPTH):
    """GPUBVHBuilder(method="median"): bvh_builder.py:488-525 with numpy's
    argpartition deciding the order inside each half (as the reference)."""
    lo, hi, cen = padded_bounds(v0, v1, v2)
    n = len(cen)
    if n == 0:
        return Tree(0)

    def leaf_test(count, depth):
        return count <= max_leaf or depth >= max_depth

    def split(start, count, depth, idx):
        bl = lo[idx].min(axis=0).tolist()
        bh = hi[idx].max(axis=0).tolist()
        axis = _longest_axis(bl, bh)
        mid = len(idx) // 2
        part = idx[np.argpartition(cen[idx, axis], mid)]
        return (start, mid, part[:mid]), (start + mid, len(idx) - mid, part[mid:])

    recs, left, right = _preorder_payload(n, leaf_test, split, np.arange(n, dtype=np.int32))
    order = np.concatenate([r[3] for r, l in zip(recs, left) if l < 0])
    return _finish(recs, left, right, order, lo, hi)


def _preorder_payload(n, leaf_test, split, payload):
    """Preorder emission: node ids in visit order, left subtree first.
    ``split(start, count, depth, payload) -> ((start, count, payload), ...)``
    gives the two children of a node that ``leaf_test(count, depth)``
    does not make a leaf."""
    recs = []
    left, right = [], []
    stack = [(0, n, 0, payload, -1, 0)]
    while stack:
        start, count, depth, pl, parent, side = stack.pop()
        idx = len(recs)
        recs.append((start, count, depth, pl))
        left.append(-1)
        right.append(-1)
        if parent >= 0:
            (left if side == 0 else right)[parent] = idx
        if not leaf_test(count, depth):
            (ls, lc, lp), (rs, rc, rp) = split(start, count, depth, pl)
            stack.append((rs, rc, depth + 1, rp, idx, 1))
            stack.append((ls, lc, depth + 1, lp, idx, 0))
    return recs, np.asarray(left, np.int32), np.asarray(right, np.int32)


def build_object_median(v0, v1, v2, max_leaf=4):
    """core/bvh.py:2-33 (BVHNode) flattened in preorder.

    Boxes come from Triangle.local_bounds (scene/mesh.py:68-76): fp64
    min/max of the vertices -/+ 1e-4.  The split sorts by fp64 centroid
    ((v0+v1)+v2)*(1/3) on the longest axis with a stable sort and cuts at
    len//2.  A node is a leaf when it holds <= max_leaf triangles or its
    depth exceeds 20; the recursion does not forward max_leaf_size
    (core/bvh.py:12-13), so only the root sees a non-default value.
    Stored boxes are the fp64 boxes rounded to fp32.
    """
    a = np.asarray(v0, np.float64).reshape(-1, 3)
    b = np.asarray(v1, np.float64).reshape(-1, 3)
    c = np.asarray(v2, np.float64).reshape(-1, 3)
    n = len(a)
    if n == 0:
        return Tree(0)
    st = np.stack([a, b, c], axis=1)
    lo = st.min(axis=1) - 1e-4
    hi = st.max(axis=1) + 1e-4
    cen = (a + b + c) * (1.0 / 3.0)  # Vec3 / s is v * (1 / s) (core/vectors.py)

    def leaf_test(count, depth):
        limit = max_leaf if depth == 0 else 4
        return count <= limit or depth > 20

    def split(start, count, depth, idx):
        axis = _longest_axis(lo[idx].min(axis=0), hi[idx].max(axis=0))
        srt = idx[np.argsort(cen[idx, axis], kind="stable")]
        mid = len(srt) // 2
        return (start, mid, srt[:mid]), (start + mid, len(srt) - mid, srt[mid:])

    recs, left, right = _preorder_payload(n, leaf_test, split, np.arange(n, dtype=np.int32))
    order = np.concatenate([r[3] for r, l in zip(recs, left) if l < 0])
    t = Tree(len(recs))
    t.left, t.right = left, right
    t.order = order.astype(np.int32)
    depth = np.asarray([r[2] for r in recs], np.int32)
    t.depth = int(depth.max())
    for i, (start, count, d, idx) in enumerate(recs):
        t.aabb_min[i] = lo[idx].min(axis=0).astype(np.float32)
        t.aabb_max[i] = hi[idx].max(axis=0).astype(np.float32)
        if left[i] < 0:
            t.tri_start[i] = start
            t.tri_count[i] = count
    return t


def material_palette(mat_rows):
    """First-appearance de-duplication of material records
    (bvh_builder.py:389-410).  ``mat_rows`` is a list of
    (type, (r,g,b), roughness, ior, emission, texture) tuples; returns the
    palette and the index of each row."""
    keys = {}
    palette = []
    idx = np.empty(len(mat_rows), np.int32)
    for i, row in enumerate(mat_rows):
        key = (int(row[0]), tuple(float(x) for x in row[1]), float(row[2]), float(row[3]), float(row[4]),
               int(row[5]))
        if key not in keys:
            keys[key] = len(palette)
            palette.append(key)
        idx[i] = keys[key]
    return palette, idx


def closed_form_shape(n, max_leaf):
    """(n_nodes, depth) of the count-median tree, by direct recursion on
    counts (memoised); used to pin the closed form."""
    from functools import lru_cache

    @lru_cache(maxsize=None)
    def rec(c):
        if c <= max_leaf:
            return 1, 0
        ln, ld = rec(c // 2)
        rn, rd = rec(c - c // 2)
        return 1 + ln + rn, 1 + max(ld, rd)

    if n == 0:
        return 0, 0
    return rec(n)
