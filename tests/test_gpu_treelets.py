"""GPU treelet restructuring (rq_bvh_treelets) against the oracle's
(or_bvh_treelets): the restructured tree node for node (child links, boxes
bit for bit, height), the canonical traversal counts on it, the answers of
every traversal variant (hit ids, t, u, v bit-exact; any hit), two rounds, a
blob round trip, and a rebuild back to the plain LBVH."""
import numpy as np
import pytest

from conftest import log
from helpers import (concat_rays, duplicate_mesh, edge_rays, mesh_configs, random_rays,
                     random_soup, rays_toward)

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _same_tree(g, o):
    for k in ("codes", "sorted_idx", "left", "right"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(_bits(g["boxes"]), _bits(o["boxes"])), "boxes"


def _answers(ctx, O, b, ob, rays, variants=False):
    o, oa = ob.closest(rays), ob.any(rays)
    g = ctx.closest_hit(b, rays, counted=True)
    for k in ("prim", "visits", "tests"):
        assert np.array_equal(g[k], o[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(_bits(g[k]), _bits(o[k])), k
    ga = ctx.any_hit(b, rays, counted=True)
    assert np.array_equal(ga["occ"], oa["occ"]) and np.array_equal(ga["visits"], oa["visits"])
    if variants:
        saved = ctx.trace_config()
        try:
            for v in ctx.trace_variants():
                ctx.trace_config(v["index"], v["index"] if not v["tstack"] else 0, 64)
                g2 = ctx.closest_hit(b, rays)
                assert np.array_equal(g2["prim"], o["prim"]), v
                assert np.array_equal(_bits(g2["t"]), _bits(o["t"])), v
                assert np.array_equal(ctx.any_hit(b, rays)["occ"], oa["occ"]), v
        finally:
            ctx.trace_config(*saved)
    return o


MESHES = dict(mesh_configs())
MESHES["dense"] = lambda: random_soup(3000, 17, spread=1.0, size=0.6)
MESHES["dup300"] = lambda: duplicate_mesh(300)


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("mode", [0, 1])
def test_treelets_match_the_oracle(ctx, O, name, mode):
    mesh = MESHES[name]()
    n = mesh.ntri
    log(f"treelets {name} mode={mode} n={n}")
    rays = concat_rays(rays_toward(3000, 71), random_rays(2000, 72), edge_rays())
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, mode)
    try:
        ob = O.OracleBVH(mesh, mode)
        ob.treelets()
        ctx.treelets(b)
        _same_tree(ctx.export(b, n), ob.arrays())
        assert ctx.bvh_info(b).height == ob.height
        _answers(ctx, O, b, ob, rays, variants=name in ("soup4097", "dense"))
        assert ctx.device_errors() == 0
    finally:
        ctx.release(b)
        ctx.release(s)


def test_treelets_large_mesh_two_rounds(ctx, O, rq):
    mesh = rq.icosphere(6, 1, center=(0, 0, -5))
    n = mesh.ntri
    rays = rays_toward(20000, 73, spread=1.2)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    try:
        ob = O.OracleBVH(mesh, 1)
        h0 = ob.height
        for r in range(2):
            log(f"treelets ico6 round {r}")
            assert ob.treelets() > 0
            ctx.treelets(b)
            _same_tree(ctx.export(b, n), ob.arrays())
        assert ctx.bvh_info(b).height == ob.height <= h0
        _answers(ctx, O, b, ob, rays, variants=True)
        # two rounds in one call: the same tree again from a fresh build
        ctx.rebuild(b)
        ctx.treelets(b, 2)
        _same_tree(ctx.export(b, n), ob.arrays())
        assert ctx.device_errors() == 0
    finally:
        ctx.release(b)
        ctx.release(s)


def test_treelets_blob_and_rebuild(ctx, O, rq):
    mesh = random_soup(20000, 23, spread=1.5, size=0.03)
    n = mesh.ntri
    rays = concat_rays(rays_toward(4000, 74), random_rays(4000, 75))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    b2 = None
    try:
        plain = O.OracleBVH(mesh, 1)
        ob = O.OracleBVH(mesh, 1)
        ob.treelets()
        ctx.treelets(b)
        blob = ctx.bvh_save(b)
        b2 = ctx.bvh_load(s, blob)
        _same_tree(ctx.export(b2, n), ob.arrays())
        _answers(ctx, O, b2, ob, rays)
        ctx.rebuild(b)
        _same_tree(ctx.export(b, n), plain.arrays())
        _answers(ctx, O, b, plain, rays)
        assert ctx.device_errors() == 0
    finally:
        if b2 is not None:
            ctx.release(b2)
        ctx.release(b)
        ctx.release(s)


def test_treelets_spheres(ctx, O, rq):
    rng = np.random.default_rng(7)
    n = 5000
    c = rng.uniform(-3, 3, (n, 3)).astype(np.float32) + np.float32([0, 0, -10])
    c[100:140] = c[99]
    r = rng.uniform(0.02, 0.5, n).astype(np.float32)
    r[100:140] = r[99]
    sph = rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(), r,
                          np.zeros(3 * n, np.float32))
    rays = random_rays(6000, 76, lo=(-4, -4, -14), hi=(4, 4, -6))
    s = ctx.upload_spheres(sph)
    b = ctx.build_lbvh(s, 1)
    try:
        ob = O.OracleBVH(sph, 1)
        ob.treelets()
        ctx.treelets(b)
        _same_tree(ctx.export(b, n), ob.arrays())
        _answers(ctx, O, b, ob, rays)
        assert ctx.device_errors() == 0
    finally:
        ctx.release(b)
        ctx.release(s)


def test_treelets_refused_rounds(ctx, rq):
    mesh = random_soup(100, 1, spread=1.5)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    try:
        for bad in (0, 9, -1):
            with pytest.raises(rq.RQError):
                ctx.treelets(b, bad)
    finally:
        ctx.release(b)
        ctx.release(s)
