"""4-wide nodes (the binary LBVH collapsed two levels at a time): the wide
nodes are exactly that collapse of the exported binary tree (slots, links,
boxes bit for bit), and every wide traversal variant returns the oracle's
answers (hit ids, t, u, v bit-exact; any hit) on triangle and sphere scenes,
after a blob load, and on a batch's BVH (which has no wide nodes and is
walked by the binary kernel)."""
import numpy as np
import pytest

import rq as _rq
from conftest import log
from helpers import (concat_rays, duplicate_mesh, edge_rays, mesh_configs, random_rays,
                     random_soup, rays_toward)

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _wide_variants(ctx):
    return [v for v in ctx.trace_variants() if v["wide"]]


def _expected_collapse(e, n, reachable_only=False):
    """Wide node of every internal node (or only of those the wide walk reaches
    from the root: the root and the even-depth nodes), from the binary export."""
    left, right, boxes = e["left"], e["right"], e["boxes"]
    empty = np.float32([np.inf, np.inf, np.inf, -np.inf, -np.inf, -np.inf])

    def box(link):
        return boxes[link] if link >= 0 else boxes[(n - 1) + ~link]

    def collapse(i):
        slots = []
        for c in (left[i], right[i]):
            if c < 0:
                slots += [(box(c), c), (empty, 0)]
            else:
                slots += [(box(left[c]), left[c]), (box(right[c]), right[c])]
        return np.stack([b for b, _ in slots]), np.int32([l for _, l in slots])

    if not reachable_only:
        return {i: collapse(i) for i in range(n - 1)}
    out = {}
    stack = [0]
    while stack:
        i = stack.pop()
        out[i] = collapse(i)
        stack += [int(l) for l in out[i][1] if l > 0]
    return out


@pytest.mark.parametrize("name", ["tri2", "dup64", "soup100", "ico2", "soup4097", "ico5"])
def test_wide_nodes_are_the_binary_collapse(ctx, rq, name):
    """Records written inside the refit chunks (from LDS) and for the nodes
    spanning chunks (from the finished binary records) alike."""
    mesh = mesh_configs()[name]()
    n = mesh.ntri
    log(f"wide collapse {name} n={n}")
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    try:
        e, w = ctx.export(b, n), ctx.export_wide(b, n)
        # the wide tree reached from the root holds every leaf exactly once
        reach = _expected_collapse(e, n, reachable_only=True)
        leaves = np.concatenate([l[l < 0] for _, l in reach.values()])
        assert np.array_equal(np.sort(~leaves), np.arange(n))
        exp = _expected_collapse(e, n)  # every internal node has its wide record
        for i, (bx, ln) in exp.items():
            assert np.array_equal(w["links"][i], ln), i
            assert np.array_equal(_bits(w["boxes"][i]), _bits(bx)), i
        assert ctx.device_errors() == 0
    finally:
        ctx.release(b)
        ctx.release(s)


def _check_variants(ctx, O, b, geo, rays, mode=1, sphere=False):
    ob = O.OracleBVH(geo, mode)
    o, oa = ob.closest(rays), ob.any(rays)
    saved = ctx.trace_config()
    try:
        for v in _wide_variants(ctx):
            for refill in (1, 16, 64):
                ctx.trace_config(v["index"], v["index"] if not v["tstack"] else 0, refill)
                g = ctx.closest_hit(b, rays)
                assert np.array_equal(g["prim"], o["prim"]), (v, refill)
                assert np.array_equal(_bits(g["t"]), _bits(o["t"])), (v, refill)
                assert np.array_equal(_bits(g["u"]), _bits(o["u"])), (v, refill)
                assert np.array_equal(_bits(g["v"]), _bits(o["v"])), (v, refill)
                ga = ctx.any_hit(b, rays)
                assert np.array_equal(ga["occ"], oa["occ"]), (v, refill)
        assert ctx.device_errors() == 0
    finally:
        ctx.trace_config(*saved)
    return o


@pytest.mark.parametrize("name", list(mesh_configs()))
def test_wide_variants_match_oracle(ctx, O, name):
    mesh = mesh_configs()[name]()
    rays = concat_rays(rays_toward(3000, 41), random_rays(2000, 42), edge_rays())
    log(f"wide variants {name}")
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    try:
        o = _check_variants(ctx, O, b, mesh, rays)
        assert (o["prim"] >= 0).any()
    finally:
        ctx.release(b)
        ctx.release(s)


def test_wide_variants_dense_overlap(ctx, O):
    """Large overlapping triangles: many slots hit per node, deep stacks."""
    mesh = random_soup(6000, 17, spread=1.0, size=0.6)
    rays = concat_rays(rays_toward(4000, 43), random_rays(4000, 44))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 0)
    try:
        _check_variants(ctx, O, b, mesh, rays, mode=0)
    finally:
        ctx.release(b)
        ctx.release(s)


def test_wide_variants_ties(ctx, O):
    mesh = duplicate_mesh(300)
    rays = concat_rays(rays_toward(2000, 45), edge_rays())
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    try:
        o = _check_variants(ctx, O, b, mesh, rays)
        assert np.all(o["prim"][o["prim"] >= 0] == 0)
    finally:
        ctx.release(b)
        ctx.release(s)


def test_wide_variants_spheres(ctx, O, rq):
    rng = np.random.default_rng(12)
    n = 4000
    c = rng.uniform(-3, 3, (n, 3)).astype(np.float32) + np.float32([0, 0, -10])
    c[100:140] = c[99]
    r = rng.uniform(0.02, 0.5, n).astype(np.float32)
    r[100:140] = r[99]
    sph = rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(), r,
                          np.zeros(3 * n, np.float32))
    d = rng.normal(size=(6000, 3))
    d[:, 2] -= 2.5
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o = np.zeros((6000, 3), np.float32)
    o[:1500] = rng.uniform(-3, 3, (1500, 3)) + np.float32([0, 0, -10])
    rays = dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(6000, np.float32),
                tmax=np.full(6000, np.inf, np.float32))
    rays["tmax"][::5] = 9.5
    rays["tmin"][::11] = 8.0
    s = ctx.upload_spheres(sph)
    b = ctx.build_lbvh(s, 1)
    try:
        out = _check_variants(ctx, O, b, sph, rays)
        assert (out["prim"] >= 0).mean() > 0.1
    finally:
        ctx.release(b)
        ctx.release(s)


def test_wide_after_blob_load(ctx, O, rq):
    """A loaded blob carries no wide nodes; the load derives them."""
    mesh = rq.icosphere(5, 2, center=(0, 0, -5))
    rays = concat_rays(rays_toward(3000, 46), random_rays(1000, 47))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    blob = ctx.bvh_save(b)
    ctx.release(b)
    b2 = ctx.bvh_load(s, blob)
    try:
        e, w = ctx.export(b2, mesh.ntri), ctx.export_wide(b2, mesh.ntri)
        for i, (bx, ln) in _expected_collapse(e, mesh.ntri).items():
            assert np.array_equal(w["links"][i], ln), i
            assert np.array_equal(_bits(w["boxes"][i]), _bits(bx)), i
        _check_variants(ctx, O, b2, mesh, rays)
    finally:
        ctx.release(b2)
        ctx.release(s)


def test_batch_bvh_wide_nodes(ctx, O, rq):
    """A batch mesh's BVH gets its 4-wide nodes from the batch kernel's widen
    phase (the binary collapse, every internal node) and again from a
    standalone rebuild of that BVH; a wide variant walks them to the
    oracle's answers."""
    mesh = random_soup(3000, 21, spread=1.5, size=0.03)
    rays = concat_rays(rays_toward(2000, 48), random_rays(1000, 49))
    s = ctx.upload_scene(mesh)
    drays = rq.DeviceRays.from_host(ctx, rays)
    n = drays.n
    hits = {f: rq.DeviceArray(ctx, 4 * n) for f in ("prim", "t", "u", "v")}
    batch = rq.Batch(ctx, [dict(scene=s, rays=drays, hits=hits, occ=None)], 1)
    saved = ctx.trace_config()
    try:
        batch.run()
        ctx.sync()
        bb = batch.bvh(0)
        o = O.OracleBVH(mesh, 1).closest(rays)
        assert np.array_equal(hits["prim"].download(np.int32, n), o["prim"])
        for rebuilt in (False, True):
            if rebuilt:
                ctx.rebuild(bb)
            e, w = ctx.export(bb, mesh.ntri), ctx.export_wide(bb, mesh.ntri)
            for i, (bx, ln) in _expected_collapse(e, mesh.ntri).items():
                assert np.array_equal(w["links"][i], ln), (rebuilt, i)
                assert np.array_equal(_bits(w["boxes"][i]), _bits(bx)), (rebuilt, i)
            for wv in (v["index"] for v in _wide_variants(ctx)):
                ctx.trace_config(wv, wv, 16)
                g = ctx.closest_hit(bb, rays)
                assert np.array_equal(g["prim"], o["prim"]), (rebuilt, wv)
                assert np.array_equal(_bits(g["t"]), _bits(o["t"])), (rebuilt, wv)
        assert ctx.device_errors() == 0
    finally:
        ctx.trace_config(*saved)
        batch.close()
        drays.free()
        for a in hits.values():
            a.free()
        ctx.release(s)


def test_wide_walk_on_a_tree_past_the_offset_limit(ctx, rq):
    """A tree of 2^25 wide nodes or more: the row offsets no longer fit 32
    bits, so the wide walks select the near / far rows per value. Their
    answers equal the binary walk's (which the oracle pins on every other
    scene) on rays through a dense soup of 2^25 + 4096 triangles."""
    n = (1 << 25) + 4096
    rng = np.random.default_rng(31)
    c = rng.random((n, 1, 3), dtype=np.float32) * np.float32(8) - np.float32(4)
    v = c + (rng.random((n, 3, 3), dtype=np.float32) - np.float32(0.5)) * np.float32(0.02)
    del c
    mesh = rq.MeshArrays(np.ascontiguousarray(v[:, :, 0].reshape(-1)),
                         np.ascontiguousarray(v[:, :, 1].reshape(-1)),
                         np.ascontiguousarray(v[:, :, 2].reshape(-1)),
                         np.arange(3 * n, dtype=np.uint32))
    del v
    rays = concat_rays(random_rays(60000, 91, lo=(-4, -4, -4), hi=(4, 4, 4)),
                       rays_toward(20000, 92, target=(0.0, 0.0, 0.0), spread=3.0,
                                   origin=(0.0, 0.0, -12.0)))
    log(f"wide walk past the offset limit: {n} triangles")
    s = ctx.upload_scene(mesh)
    del mesh
    b = ctx.build_lbvh(s, 1)
    saved = ctx.trace_config()
    try:
        assert ctx.bvh_info(b).n_internal >= 1 << 25
        ctx.trace_config(0, 0, 64)
        o, oa = ctx.closest_hit(b, rays), ctx.any_hit(b, rays)
        assert (o["prim"] >= 0).mean() > 0.3
        for wv in _wide_variants(ctx):
            for refill in (16, 64):
                ctx.trace_config(wv["index"], wv["index"] if not wv["tstack"] else 0, refill)
                g = ctx.closest_hit(b, rays)
                assert np.array_equal(g["prim"], o["prim"]), (wv, refill)
                assert np.array_equal(_bits(g["t"]), _bits(o["t"])), (wv, refill)
                assert np.array_equal(ctx.any_hit(b, rays)["occ"], oa["occ"]), (wv, refill)
        assert ctx.device_errors() == 0
    finally:
        ctx.trace_config(*saved)
        ctx.release(b)
        ctx.release(s)
