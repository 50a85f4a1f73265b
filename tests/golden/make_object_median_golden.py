"""Golden vectors of core/bvh.py's BVHNode, from the reference's own code.

core/bvh.py imports nothing outside the reference's ``core`` package (and
that package only the standard library), so it is imported here, by path,
and run on small seeded meshes.  Its triangles come from scene/mesh.py,
which imports taichi and cannot be loaded here; ``_Tri`` stands in for
them with the two methods BVHNode calls, local_bounds (scene/mesh.py:68-76)
and centroid (:78-79), computed on the reference's own Vec3 and AABB.

The tree is flattened in preorder (node, left subtree, right subtree):
fp64 boxes, child ids (-1 at leaves), leaf ranges and the leaf order of the
triangle indices, stored per case in

    tests/golden/object_median.npz  {case<k>_<field>, cases}

Run from the repository root, in a container that holds the reference:
    python tests/golden/make_object_median_golden.py [/path/to/reference]
The tests read only the .npz.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
REFERENCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
sys.path.insert(0, str(REFERENCE))

from core.aabb import AABB  # noqa: E402
from core.bvh import BVHNode  # noqa: E402
from core.vectors import Vec3  # noqa: E402


class _Tri:
    """The part of scene/mesh.py's Triangle that BVHNode uses."""

    def __init__(self, index, a, b, c):
        self.index = index
        self._v0, self._v1, self._v2 = Vec3(*a), Vec3(*b), Vec3(*c)

    def local_bounds(self):
        eps = 1e-4
        lo = Vec3(min(self._v0.x, self._v1.x, self._v2.x) - eps, min(self._v0.y, self._v1.y, self._v2.y) - eps,
                  min(self._v0.z, self._v1.z, self._v2.z) - eps)
        hi = Vec3(max(self._v0.x, self._v1.x, self._v2.x) + eps, max(self._v0.y, self._v1.y, self._v2.y) + eps,
                  max(self._v0.z, self._v1.z, self._v2.z) + eps)
        return AABB(lo, hi)

    def centroid(self):
        return (self._v0 + self._v1 + self._v2) / 3.0


def _mesh(n, seed, kind):
    rng = np.random.default_rng(seed)
    if kind == "random":
        v0 = rng.normal(size=(n, 3)) * 10.0
        return v0, v0 + rng.normal(size=(n, 3)) * 0.3, v0 + rng.normal(size=(n, 3)) * 0.3
    if kind == "grid":
        v0 = rng.integers(0, 6, size=(n, 3)).astype(np.float64)
        return v0, v0 + [1.0, 0.0, 0.0], v0 + [0.0, 1.0, 0.0]
    if kind == "flat":
        v0 = np.stack([rng.integers(0, 16, n), rng.integers(0, 16, n), np.zeros(n)], 1).astype(np.float64)
        return v0, v0 + [1.0, 0.0, 0.0], v0 + [0.0, 1.0, 0.0]
    v0 = np.where(rng.random((n, 3)) < 0.5, -0.0, 0.0)
    v0[: n // 4] = rng.normal(size=(n // 4, 3))
    return v0, v0.copy(), v0.copy()


CASES = [(5, 4, "random", 1), (37, 16, "random", 2), (300, 4, "grid", 3), (200, 1, "flat", 4), (64, 4, "zeros", 5),
         (2000, 4, "random", 6)]


def _flatten(root):
    lo, hi, left, right, start, count, order = [], [], [], [], [], [], []

    def visit(node):
        me = len(lo)
        lo.append([node.bounds.min.x, node.bounds.min.y, node.bounds.min.z])
        hi.append([node.bounds.max.x, node.bounds.max.y, node.bounds.max.z])
        left.append(-1)
        right.append(-1)
        start.append(-1)
        count.append(0)
        if node.triangles is not None:
            start[me] = len(order)
            count[me] = len(node.triangles)
            order.extend(t.index for t in node.triangles)
        else:
            left[me] = visit(node.left)
            right[me] = visit(node.right)
        return me

    visit(root)
    return {"aabb_min": np.asarray(lo, np.float64), "aabb_max": np.asarray(hi, np.float64),
            "left": np.asarray(left, np.int32), "right": np.asarray(right, np.int32),
            "tri_start": np.asarray(start, np.int32), "tri_count": np.asarray(count, np.int32),
            "order": np.asarray(order, np.int32)}


def main():
    out = {"cases": np.asarray([[n, leaf, seed] for n, leaf, _, seed in CASES], np.int64)}
    for k, (n, leaf, kind, seed) in enumerate(CASES):
        v0, v1, v2 = _mesh(n, seed, kind)
        tris = [_Tri(i, v0[i].tolist(), v1[i].tolist(), v2[i].tolist()) for i in range(n)]
        tree = _flatten(BVHNode(tris, max_leaf_size=leaf))
        for name, arr in (("v0", v0), ("v1", v1), ("v2", v2), *tree.items()):
            out[f"case{k}_{name}"] = arr
        print(f"case {k}: n={n} root leaf={leaf} {kind}: {len(tree['left'])} nodes")
    np.savez_compressed(ROOT / "golden" / "object_median.npz", **out)


if __name__ == "__main__":
    main()
