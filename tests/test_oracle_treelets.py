"""CPU: the oracle's treelet restructuring (Karras & Aila 2013, one round of
7-leaf treelets). The subset dynamic programme is pinned by brute force over
every binary tree of 2..7 leaves (float32 costs, the minimum exactly); the
restructured trees keep their leaf set, their boxes are the unions of their
children bit for bit, their cost and height never grow, and queries return
exactly what they returned on the plain LBVH (and brute force)."""
import itertools

import numpy as np
import pytest

from helpers import (concat_rays, duplicate_mesh, edge_rays, mesh_configs, random_rays,
                     random_soup, rays_toward)


def _half_area(b):
    b = np.asarray(b, np.float32)
    dx, dy, dz = b[3] - b[0], b[4] - b[1], b[5] - b[2]
    return np.float32(np.float32(dx * dy + dy * dz) + dz * dx)


def _all_trees(boxes):
    """Every full binary tree over the boxes: cost (float32, internal half areas
    summed as node + (left + right)) and height, per subset, brute force."""
    m = len(boxes)
    memo = {}

    def union(s):
        sel = boxes[[k for k in range(m) if s >> k & 1]]
        return np.concatenate([sel[:, :3].min(0), sel[:, 3:].max(0)]).astype(np.float32)

    def trees(s):
        if s in memo:
            return memo[s]
        if s & (s - 1) == 0:
            memo[s] = [(np.float32(0), 0)]
            return memo[s]
        a = _half_area(union(s))
        low = s & -s
        rest = s ^ low
        out = []
        bits = [1 << k for k in range(m) if rest >> k & 1]
        for r in range(len(bits)):
            for comb in itertools.combinations(bits, r):
                p = low | sum(comb)
                for cl, hl in trees(p):
                    for cr, hr in trees(s ^ p):
                        out.append((np.float32(a + np.float32(cl + cr)), 1 + max(hl, hr)))
        memo[s] = out
        return out

    return trees((1 << m) - 1)


def _boxes(rng, m, kind):
    lo = rng.uniform(-1, 1, (m, 3)).astype(np.float32)
    ext = rng.uniform(0, 0.8, (m, 3)).astype(np.float32)
    if kind == "flat":
        ext[:, 2] = 0
    if kind == "same":
        lo[:] = lo[0]
        ext[:] = ext[0]
    if kind == "pairs":
        lo[1::2] = lo[0::2][: len(lo[1::2])]
        ext[1::2] = ext[0::2][: len(ext[1::2])]
    return np.concatenate([lo, lo + ext], 1).astype(np.float32)


@pytest.mark.parametrize("m", [2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("kind", ["random", "flat", "same", "pairs"])
def test_treelet_dp_is_the_brute_force_minimum(O, m, kind):
    rng = np.random.default_rng(100 * m + len(kind))
    for _ in range(3 if m == 7 else 6):
        b = _boxes(rng, m, kind)
        cost, split, height = O.treelet_dp(b)
        trees = _all_trees(b)
        assert len(trees) == int(np.prod(np.arange(1, 2 * m - 2, 2)))  # (2m-3)!!
        best = min(c for c, _ in trees)
        assert np.float32(cost) == best
        # the split table describes a tree of that cost and of its stated height
        def walk(s):
            if s & (s - 1) == 0:
                return np.float32(0), 0
            p = int(split[s])
            assert p & s == p and p & (s & -s)
            cl, hl = walk(p)
            cr, hr = walk(s ^ p)
            sel = b[[k for k in range(m) if s >> k & 1]]
            u = np.concatenate([sel[:, :3].min(0), sel[:, 3:].max(0)])
            return np.float32(_half_area(u) + np.float32(cl + cr)), 1 + max(hl, hr)
        c, h = walk((1 << m) - 1)
        assert c == best and h == height[(1 << m) - 1]


def _cost(boxes, n):
    b = boxes[: n - 1].astype(np.float64)
    d = b[:, 3:] - b[:, :3]
    return float((d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]).sum())


def _check_tree(a, n):
    left, right, boxes = a["left"], a["right"], a["boxes"]
    seen_int = np.zeros(n - 1, np.int32)
    seen_leaf = np.zeros(n, np.int32)
    stack = [0]
    while stack:
        i = stack.pop()
        seen_int[i] += 1
        for c in (left[i], right[i]):
            if c >= 0:
                stack.append(int(c))
            else:
                seen_leaf[~c] += 1
    assert np.all(seen_int == 1) and np.all(seen_leaf == 1)

    def box(c):
        return boxes[c] if c >= 0 else boxes[(n - 1) + ~c]

    for i in range(n - 1):
        bl, br = box(left[i]), box(right[i])
        u = np.concatenate([np.minimum(bl[:3], br[:3]), np.maximum(bl[3:], br[3:])])
        assert np.array_equal(u.view(np.uint32), boxes[i].view(np.uint32)), i


def _tree_height(a, n):
    best, stack = 0, [(0, 0)]
    while stack:
        i, d = stack.pop()
        for c in (a["left"][i], a["right"][i]):
            if c >= 0:
                stack.append((int(c), d + 1))
            else:
                best = max(best, d + 1)
    return best


MESHES = dict(mesh_configs())
MESHES["dense"] = lambda: random_soup(3000, 17, spread=1.0, size=0.6)
MESHES["dup300"] = lambda: duplicate_mesh(300)


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("mode", [0, 1])
def test_treelets_keep_the_tree_valid_and_the_answers(O, name, mode):
    mesh = MESHES[name]()
    n = mesh.ntri
    ob = O.OracleBVH(mesh, mode)
    before = ob.arrays()
    rays = concat_rays(rays_toward(3000, 61), random_rays(2000, 62), edge_rays())
    o0, a0 = ob.closest(rays), ob.any(rays)
    h0 = ob.height
    replaced = ob.treelets()
    after = ob.arrays()
    assert np.array_equal(after["sorted_idx"], before["sorted_idx"])
    assert np.array_equal(after["codes"], before["codes"])
    if n < 7:
        assert replaced == 0
        return
    _check_tree(after, n)
    # leaf boxes unchanged, root box unchanged
    assert np.array_equal(after["boxes"][n - 1:].view(np.uint32),
                          before["boxes"][n - 1:].view(np.uint32))
    assert np.array_equal(after["boxes"][0].view(np.uint32), before["boxes"][0].view(np.uint32))
    assert _cost(after["boxes"], n) <= _cost(before["boxes"], n) * (1 + 1e-6)
    assert ob.height == _tree_height(after, n) <= h0
    o1, a1 = ob.closest(rays), ob.any(rays)
    for k in ("prim", "t", "u", "v"):
        assert np.array_equal(np.asarray(o1[k]).view(np.uint32),
                              np.asarray(o0[k]).view(np.uint32)), k
    assert np.array_equal(a1["occ"], a0["occ"])
    if n <= 5000:
        bf = O.closest_brute(mesh, rays)
        assert np.array_equal(o1["prim"], bf["prim"])


def test_treelets_improve_a_large_mesh(O, rq):
    mesh = rq.icosphere(6, 1, center=(0, 0, -5))
    ob = O.OracleBVH(mesh, 1)
    c0 = _cost(ob.arrays()["boxes"], mesh.ntri)
    rays = rays_toward(20000, 63, spread=1.2)
    v0 = ob.closest(rays)["visits"].sum()
    replaced = ob.treelets()
    assert replaced > mesh.ntri // 20
    assert _cost(ob.arrays()["boxes"], mesh.ntri) < 0.95 * c0
    assert ob.closest(rays)["visits"].sum() < v0


def test_treelets_spheres(O, rq):
    rng = np.random.default_rng(5)
    n = 3000
    c = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    r = rng.uniform(0.02, 0.4, n).astype(np.float32)
    sph = rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(), r,
                          np.zeros(3 * n, np.float32))
    ob = O.OracleBVH(sph, 1)
    rays = random_rays(4000, 64, lo=(-4, -4, -4), hi=(4, 4, 4))
    o0 = ob.closest(rays)
    assert ob.treelets() > 0
    _check_tree(ob.arrays(), n)
    o1 = ob.closest(rays)
    assert np.array_equal(o1["prim"], o0["prim"])
    assert np.array_equal(o1["t"].view(np.uint32), o0["t"].view(np.uint32))
