"""CPU: the oracle's sort, Karras topology and refit against independent
algorithms (numpy lexsort, a top-down split builder, a linear-scan split
search) and structural invariants."""
import numpy as np
import pytest

from helpers import duplicate_mesh, random_soup


def _brute_topology(sc: np.ndarray):
    """Top-down radix tree by linear scan over 64-bit keys (code<<32 | i)."""
    n = len(sc)
    keys = [(int(c) << 32) | i for i, c in enumerate(sc)]

    def prefix(a, b):
        x = keys[a] ^ keys[b]
        return 64 - x.bit_length()

    left = [0] * (n - 1)
    right = [0] * (n - 1)
    stack = [(0, n - 1, 0)]
    while stack:
        a, b, node = stack.pop()
        p = prefix(a, b)
        split = max(s for s in range(a, b) if prefix(a, s) > p)
        left[node] = ~split if split == a else split
        right[node] = ~(split + 1) if split + 1 == b else split + 1
        if split != a:
            stack.append((a, split, split))
        if split + 1 != b:
            stack.append((split + 1, b, split + 1))
    return np.array(left, np.int32), np.array(right, np.int32)


@pytest.mark.parametrize("seed,n,span", [(1, 2, 4), (2, 3, 4), (3, 17, 4), (4, 64, 16),
                                         (5, 100, 1 << 30), (6, 257, 8), (7, 50, 1)])
def test_karras_three_ways(O, seed, n, span):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, span, n).astype(np.uint32)
    sc, si = O.sort_pairs(codes)
    l1, r1 = O.karras(sc)
    l2, r2 = O.topdown(sc)
    l3, r3 = _brute_topology(sc)
    assert np.array_equal(l1, l2) and np.array_equal(r1, r2)
    assert np.array_equal(l1, l3) and np.array_equal(r1, r3)


def test_sort_is_stable_lexsort(O):
    rng = np.random.default_rng(11)
    codes = rng.integers(0, 50, 100000).astype(np.uint32)
    sc, si = O.sort_pairs(codes)
    order = np.lexsort((np.arange(codes.size), codes))
    assert np.array_equal(si, order)
    assert np.array_equal(sc, codes[order])


def _ranges(left, right, n):
    first = np.zeros(n - 1, np.int64)
    last = np.zeros(n - 1, np.int64)
    parent_count = np.zeros(2 * n - 1, np.int64)
    order, st = [], [0]
    while st:
        i = st.pop()
        order.append(i)
        for c in (left[i], right[i]):
            parent_count[c if c >= 0 else (n - 1) + ~c] += 1
            if c >= 0:
                st.append(c)
    for i in reversed(order):
        l, r = left[i], right[i]
        first[i] = ~l if l < 0 else first[l]
        last[i] = ~r if r < 0 else last[r]
    return first, last, parent_count, len(order)


@pytest.mark.parametrize("mk", [lambda rq: rq.icosphere(4, 3, center=(0, 0, -5)),
                                lambda rq: random_soup(3000, 8),
                                lambda rq: duplicate_mesh(777)])
def test_topology_invariants(O, rq, mk):
    mesh = mk(rq)
    b = O.OracleBVH(mesh, 1)
    a = b.arrays()
    n = mesh.ntri
    first, last, pc, reached = _ranges(a["left"], a["right"], n)
    assert reached == n - 1                       # every internal node reachable
    assert (pc[1:] == 1).all() and pc[0] == 0     # one parent each, root has none
    assert first[0] == 0 and last[0] == n - 1     # root covers everything
    idx = np.arange(n - 1)
    assert ((first == idx) | (last == idx)).all()  # Karras: node i is an end of its range
    assert b.height <= 63


def test_refit_boxes(O, rq):
    mesh = rq.icosphere(3, 9, center=(1, 2, -5))
    b = O.OracleBVH(mesh, 1)
    a = b.arrays()
    n = mesh.ntri
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1)
    tri = v[mesh.idx.reshape(-1, 3)][a["sorted_idx"]]
    leaf = a["boxes"][n - 1:]
    assert np.array_equal(leaf[:, :3], tri.min(axis=1) + 0.0)
    assert np.array_equal(leaf[:, 3:], tri.max(axis=1) + 0.0)

    def box(c):
        return a["boxes"][c] if c >= 0 else a["boxes"][(n - 1) + ~c]

    for i in range(n - 1):
        l, r = box(a["left"][i]), box(a["right"][i])
        assert np.array_equal(a["boxes"][i][:3], np.minimum(l[:3], r[:3]))
        assert np.array_equal(a["boxes"][i][3:], np.maximum(l[3:], r[3:]))
    assert np.array_equal(a["boxes"][0][:3], v.min(axis=0))
    assert np.array_equal(a["boxes"][0][3:], v.max(axis=0))
