"""Cached BVH blob: a saved and reloaded LBVH (no rebuild) is the same tree
(codes, order, topology, boxes, height) and answers every query like the
original; damaged or foreign blobs are refused."""
import numpy as np
import pytest

from conftest import log
from helpers import random_rays, random_soup, rays_toward

pytestmark = pytest.mark.gpu


def _same_tree(ctx, a, b, n):
    ea, eb = ctx.export(a, n), ctx.export(b, n)
    for k in ("codes", "sorted_idx", "left", "right"):
        assert np.array_equal(ea[k], eb[k]), k
    assert np.array_equal(ea["boxes"].view(np.uint32), eb["boxes"].view(np.uint32))
    ia, ib = ctx.bvh_info(a), ctx.bvh_info(b)
    assert ia.height == ib.height and list(ia.root_lo) == list(ib.root_lo)


def _same_answers(ctx, a, b, rays):
    ca, cb = ctx.closest_hit(a, rays, counted=True), ctx.closest_hit(b, rays, counted=True)
    for k in ("prim", "visits", "tests"):
        assert np.array_equal(ca[k], cb[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(ca[k].view(np.uint32), cb[k].view(np.uint32)), k
    assert np.array_equal(ctx.any_hit(a, rays)["occ"], ctx.any_hit(b, rays)["occ"])


@pytest.mark.parametrize("which", ["ico5", "soup", "tri1", "spheres"])
def test_blob_roundtrip(ctx, rq, tmp_path, which):
    log(f"blob {which}")
    if which == "spheres":
        rng = np.random.default_rng(3)
        c = rng.uniform(-2, 2, (3000, 3)).astype(np.float32) + np.float32([0, 0, -5])
        sph = rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(),
                              rng.uniform(0.02, 0.2, 3000).astype(np.float32),
                              np.zeros(9000, np.float32))
        s, n = ctx.upload_spheres(sph), 3000
    else:
        mesh = {"ico5": lambda: rq.icosphere(5, 2, center=(0, 0, -5)),
                "soup": lambda: random_soup(9001, 4, spread=1.5, size=0.03),
                "tri1": lambda: random_soup(1, 5, spread=0.2, size=1.0)}[which]()
        s, n = ctx.upload_scene(mesh), mesh.ntri
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    blob = ctx.bvh_save(b)
    b2 = ctx.bvh_load(s, blob)
    path = tmp_path / "t.rqbvh"
    ctx.bvh_save_file(b, path)
    b3 = ctx.bvh_load_file(s, path)
    try:
        assert path.read_bytes() == blob
        rays = dict(rays_toward(4000, 1)) if which != "spheres" else None
        if rays is None:
            from helpers import concat_rays
            rays = concat_rays(rays_toward(3000, 2), random_rays(2000, 3))
        for other in (b2, b3):
            _same_tree(ctx, b, other, n)
            _same_answers(ctx, b, other, rays)
        ctx.rebuild(b2)  # a loaded BVH rebuilds from its scene like any other
        _same_tree(ctx, b, b2, n)
    finally:
        for h in (b3, b2, b):
            ctx.release(h)
        ctx.release(s)


def test_blob_refusals(ctx, rq):
    mesh = rq.icosphere(3, 7, center=(0, 0, -5))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    blob = bytearray(ctx.bvh_save(b))
    try:
        for damage in ("payload", "magic", "truncate"):
            bad = bytearray(blob)
            if damage == "payload":
                bad[200] ^= 0x40
            elif damage == "magic":
                bad[0] ^= 0x01
            else:
                bad = bad[: len(bad) // 2]
            with pytest.raises(rq.RQError) as e:
                ctx.bvh_load(s, bytes(bad))
            assert e.value.code == -4, damage  # RQ_E_IO
        # same triangle count, other vertices: the fingerprint refuses it
        moved = rq.MeshArrays(mesh.vx + np.float32(0.5), mesh.vy, mesh.vz, mesh.idx)
        s2 = ctx.upload_scene(moved)
        with pytest.raises(rq.RQError) as e:
            ctx.bvh_load(s2, bytes(blob))
        assert e.value.code == -1  # RQ_E_INVALID
        ctx.release(s2)
    finally:
        ctx.release(b)
        ctx.release(s)


@pytest.mark.parametrize("which", ["ico5", "spheres", "tri1"])
def test_saved_blob_equals_the_oracle_tree(ctx, rq, O, which):
    """rq_bvh_save writes, byte for byte, the blob of the oracle's tree
    (helpers.blob_payload: codes, order, parent links, node records, leaf
    records, root box, height, FNV-1a 64 checksum and scene fingerprint)."""
    from helpers import blob_from_payload, blob_payload

    log(f"blob vs oracle {which}")
    if which == "spheres":
        rng = np.random.default_rng(9)
        c = rng.uniform(-2, 2, (2000, 3)).astype(np.float32)
        prims = rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(),
                                rng.uniform(0.02, 0.2, 2000).astype(np.float32),
                                np.zeros(6000, np.float32))
        s = ctx.upload_spheres(prims)
    else:
        prims = (rq.icosphere(5, 2, center=(0, 0, -5)) if which == "ico5"
                 else random_soup(1, 5, spread=0.2, size=1.0))
        s = ctx.upload_scene(prims)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    try:
        blob = ctx.bvh_save(b)
        ob = O.OracleBVH(prims, rq.MORTON_SCENE_AABB)
        mine = blob_from_payload(prims, blob_payload(prims, ob.arrays(), ob.height))
        assert blob == mine
        assert rq.bvh_blob_validate(blob, prims)[0] == rq.RQ_OK
    finally:
        ctx.release(b)
        ctx.release(s)


def test_crafted_blob_refused_before_upload(ctx, rq):
    """A blob whose checksum was recomputed after an edit (a child link out
    of range) is refused by the host check of rq_bvh_load (RQ_E_IO); the
    device never sees it."""
    from helpers import fnv1a64

    mesh = rq.icosphere(3, 7, center=(0, 0, -5))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    try:
        blob = bytearray(ctx.bvh_save(b))
        n = mesh.ntri
        link0 = 64 + 4 * n + 4 * n + 4 * (2 * n - 1) + 12  # node 0, left link
        blob[link0:link0 + 4] = np.int32(5 * n).tobytes()
        blob[48:56] = np.uint64(fnv1a64(bytes(blob[64:]))).tobytes()
        with pytest.raises(rq.RQError) as e:
            ctx.bvh_load(s, bytes(blob))
        assert e.value.code == rq.RQ_E_IO
        assert ctx.device_errors() == 0
    finally:
        ctx.release(b)
        ctx.release(s)
