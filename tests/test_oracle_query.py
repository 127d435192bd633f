"""CPU: the oracle's Möller–Trumbore test (closed-form known answers), its
canonical BVH traversal against brute force, the reference's own sphere
test, and the CPU baseline port."""
import numpy as np
import pytest

from helpers import (RAY_FIELDS, concat_rays, duplicate_mesh, edge_rays, mesh_configs,
                     random_rays, random_soup, rays_toward)

V0, V1, V2 = (0.0, 0.0, -5.0), (1.0, 0.0, -5.0), (0.0, 1.0, -5.0)


@pytest.mark.parametrize("o,d,expect", [
    ((0.25, 0.25, 0.0), (0, 0, -1), (True, 5.0, 0.25, 0.25)),
    ((0.5, 0.0, 0.0), (0, 0, -1), (True, 5.0, 0.5, 0.0)),     # on an edge: v = 0 counts
    ((0.0, 0.0, 0.0), (0, 0, -1), (True, 5.0, 0.0, 0.0)),     # on a vertex
    ((0.75, 0.75, 0.0), (0, 0, -1), (False,)),                # u + v > 1
    ((-0.25, 0.25, 0.0), (0, 0, -1), (False,)),               # u < 0
    ((0.25, 0.25, 0.0), (1, 0, 0), (False,)),                 # parallel: |det| < EPS
    ((0.25, 0.25, -10.0), (0, 0, 1), (True, 5.0, 0.25, 0.25)),  # from behind (no culling)
    ((0.25, 0.25, 0.0), (0, 0, 1), (False,)),                 # t = -5 is not in (0, inf)
])
def test_mt_known_answers(O, o, d, expect):
    hit, t, u, v = O.mt(V0, V1, V2, o, d)
    assert hit == expect[0]
    if hit:
        assert (t, u, v) == expect[1:]


def test_mt_interval_is_open(O):
    o, d = (0.25, 0.25, 0.0), (0, 0, -1)
    assert not O.mt(V0, V1, V2, o, d, 0.0, 5.0)[0]   # t == tmax excluded
    assert not O.mt(V0, V1, V2, o, d, 5.0, 9.0)[0]   # t == tmin excluded
    assert O.mt(V0, V1, V2, o, d, 4.999, 5.001)[0]


def test_mt_tiny_determinant(O):
    s = 1e-4  # edges 1e-4: det = 1e-8 < EPS (main.cpp:59) -> rejected
    hit = O.mt((0, 0, -5), (s, 0, -5), (0, s, -5), (s / 4, s / 4, 0), (0, 0, -1))[0]
    assert not hit


@pytest.mark.parametrize("name", list(mesh_configs()))
@pytest.mark.parametrize("mode", [0, 1])
def test_bvh_equals_brute_force(O, name, mode):
    mesh = mesh_configs()[name]()
    rays = concat_rays(rays_toward(3000, 11), random_rays(2000, 12), edge_rays())
    b = O.OracleBVH(mesh, mode)
    o = b.closest(rays)
    bf = O.closest_brute(mesh, rays)
    assert np.array_equal(o["prim"], bf["prim"])
    assert np.array_equal(o["t"].view(np.uint32), bf["t"].view(np.uint32))
    assert np.array_equal(o["u"].view(np.uint32), bf["u"].view(np.uint32))
    assert np.array_equal(b.any(rays)["occ"], O.any_brute(mesh, rays))


def test_bvh_equals_brute_force_vertex_aimed(O, rq):
    """Rays aimed exactly at mesh vertices (hits on edges and vertices)."""
    mesh = rq.icosphere(5, 2, center=(0, 0, -5))
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1)
    rng = np.random.default_rng(0)
    rows = []
    for org in [(0, 0, 0), (0, 0, -5), (0.3, -0.2, 1.0)]:
        o = np.array(org, np.float32)
        d = (v[rng.choice(len(v), 4000)] - o).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rows.append(np.column_stack([np.tile(o, (len(d), 1)), d, np.zeros(len(d)),
                                     np.full(len(d), np.inf)]).astype(np.float32))
    a = np.concatenate(rows)
    rays = {k: a[:, i].copy() for i, k in enumerate(RAY_FIELDS)}
    b = O.OracleBVH(mesh, 1)
    o = b.closest(rays)
    bf = O.closest_brute(mesh, rays)
    assert np.array_equal(o["prim"], bf["prim"])
    assert np.array_equal(o["t"].view(np.uint32), bf["t"].view(np.uint32))


def test_ties_lowest_index(O):
    mesh = duplicate_mesh(50)
    rays = rays_toward(500, 5, spread=0.5)
    o = O.OracleBVH(mesh, 1).closest(rays)
    assert (o["prim"][o["prim"] >= 0] == 0).all() and (o["prim"] >= 0).sum() > 100


def test_canonical_counts(O, rq):
    """Visits count two node records per internal node; a miss of the root's
    children costs exactly 2 visits and no test."""
    mesh = rq.icosphere(2, 1, center=(0, 0, -5))
    rays = rays_toward(1, 0, target=(0, 0, 50), spread=0.0)  # points away from the mesh
    o = O.OracleBVH(mesh, 1).closest(rays)
    assert o["prim"][0] == -1 and o["visits"][0] == 2 and o["tests"][0] == 0
    o = O.OracleBVH(mesh, 1).closest(rays_toward(500, 1))
    assert (o["visits"] % 2 == 0).all() and (o["tests"] >= (o["prim"] >= 0)).all()


def test_sphere_test_matches_reference(O):
    ref = O.ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    import ctypes as C

    rng = np.random.default_rng(3)
    for _ in range(3000):
        c = rng.uniform(-5, 5, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.05, 3))
        o = rng.uniform(-5, 5, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        a0, a1, b0, b1 = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        h1 = O.lib().or_ray_sphere(O._p(c), float(r), O._p(o), O._p(d), C.byref(a0),
                                   C.byref(a1))
        h2 = ref.ref_ray_sphere(O._p(c), float(r), O._p(o), O._p(d), C.byref(b0), C.byref(b1))
        assert h1 == h2
        if h1:
            assert (a0.value, a1.value) == (b0.value, b1.value)


def test_port_baseline_matches_brute_force(O):
    """The CPU baseline (reference median-split BVH + collect-then-test) finds
    the same closest triangle wherever the closest t is unique."""
    mesh = random_soup(3000, 21, spread=1.0, size=0.05)
    rays = random_rays(3000, 22, lo=(-1.2, -1.2, -1.2), hi=(1.2, 1.2, 1.2))
    p = O.PortBVH(mesh).closest(rays, threads=4)
    bf = O.closest_brute(mesh, rays)
    assert np.array_equal(p["prim"], bf["prim"])
    assert O.PortBVH(mesh).n_nodes == 2 * 3000 - 1


@pytest.mark.parametrize("ntri", [300, 3000])
def test_port_reference_lbvh_matches_brute_force(O, ntri):
    """The restated constructLBVHTree (Morton codes, radixSort of a copy,
    root split, halves on two threads above 500 objects) builds a tree whose
    collect-then-test closest hit equals brute force, every object kept."""
    mesh = random_soup(ntri, 23, spread=1.0, size=0.05)
    rays = random_rays(3000, 24, lo=(-1.2, -1.2, -1.2), hi=(1.2, 1.2, 1.2))
    t = O.PortBVH(mesh, lbvh=True)
    assert t.n_nodes == 2 * ntri - 1
    p = t.closest(rays, threads=4)
    assert np.array_equal(p["prim"], O.closest_brute(mesh, rays)["prim"])


def _slab_hits(boxes, rays, slack=np.float32(1.000030517578125)):
    """Boxes (m, 6) hit by each ray with tmax = +inf, by the conservative slab
    test (near planes by the sign of 1/d, far bound times the slack)."""
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1).astype(np.float32)
    with np.errstate(divide="ignore"):
        inv = (np.float32(1) / np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)).astype(
            np.float32)
    lo, hi = boxes[None, :, :3], boxes[None, :, 3:]
    neg = np.signbit(inv)[:, None, :]
    near = np.where(neg, hi, lo)
    far = np.where(neg, lo, hi)
    tn = ((near - o[:, None, :]) * inv[:, None, :]).astype(np.float32)
    tf = ((far - o[:, None, :]) * inv[:, None, :]).astype(np.float32)
    tmin = rays["tmin"][:, None].astype(np.float32)
    tnear = np.fmax(np.fmax(tn[..., 0], tn[..., 1]), np.fmax(tn[..., 2], tmin))
    tfar = np.fmin(np.fmin(tf[..., 0], tf[..., 1]), tf[..., 2])
    return tnear <= (tfar * slack).astype(np.float32)


def test_counts_of_rays_that_hit_nothing(O):
    """For a ray that hits no triangle nothing is culled, and child boxes lie
    inside their parents' (exact unions, monotone slab test), so the walk
    reaches exactly the nodes whose own box the ray hits: visits = 2 x (the
    root + internal nodes with a hit box), tests = leaves with a hit box."""
    mesh = random_soup(300, 31, spread=1.0, size=0.08)
    rays = random_rays(3000, 32, lo=(-1.3, -1.3, -1.3), hi=(1.3, 1.3, 1.3))
    ob = O.OracleBVH(mesh, 1)
    o = ob.closest(rays)
    a = ob.any(rays)
    miss = np.flatnonzero(o["prim"] < 0)
    assert len(miss) > 1000
    sub = {k: v[miss] for k, v in rays.items()}
    boxes = ob.arrays()["boxes"].astype(np.float32)
    n = mesh.ntri
    hit = _slab_hits(boxes, sub)
    internal = hit[:, 1:n - 1].sum(1)   # rows 1..n-2: internal nodes other than the root
    leaves = hit[:, n - 1:].sum(1)      # rows n-1..: leaves
    assert np.array_equal(o["visits"][miss], 2 * (1 + internal))
    assert np.array_equal(o["tests"][miss], leaves)
    assert np.array_equal(a["visits"][miss], o["visits"][miss])
    assert np.array_equal(a["tests"][miss], o["tests"][miss])
    assert leaves.max() > 3


def test_counts_when_every_box_is_hit(O, rq):
    """n identical triangles and a ray through them all at one t: closest hit
    visits every internal node (2 (n - 1) records) and tests every leaf; any
    hit follows the left links (ties go left) to the first leaf and stops."""
    for n in (2, 37, 64):
        mesh = duplicate_mesh(n)
        ob = O.OracleBVH(mesh, 1)
        rays = rays_toward(1, 0, target=(0.0, 0.0, -5.0), spread=0.0)
        o, a = ob.closest(rays), ob.any(rays)
        assert o["prim"][0] == 0 and o["visits"][0] == 2 * (n - 1) and o["tests"][0] == n
        left = ob.arrays()["left"]
        depth, node = 0, 0
        while node >= 0:
            depth += 1
            node = left[node]
        assert a["occ"][0] == 1 and a["tests"][0] == 1 and a["visits"][0] == 2 * depth
