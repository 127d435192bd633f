"""The allocation-free closest-hit entry (rq_closest_hit_noalloc), called
through the C ABI, against rq_closest_hit: identical prim, t, u, v on two
calls in a row (the second smaller, so stale scratch contents of the first
would show), on a stream that never queried (rq_closest_hit would allocate
its overflow area there), with no change in the context's device bytes; the
first call in a variant with 12 LDS stack entries over a scene whose walks
need more than 16, so the scratch's overflow area is used."""
import ctypes as C

import numpy as np
import pytest

import rq as _rq
from conftest import log
from helpers import concat_rays, random_rays, random_soup, rays_toward

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _noalloc(ctx, bvh, rays, hits, stream):
    lib = _rq.lib()
    fn = lib.rq_closest_hit_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_rq.Rays), C.c_uint32,
                   C.POINTER(_rq.Hits), C.c_void_p]
    r = rays.struct()
    h = _rq.Hits(hits["prim"].ptr, hits["t"].ptr, hits["u"].ptr, hits["v"].ptr)
    ctx._check(fn(ctx.h, bvh, C.byref(r), rays.n, C.byref(h), stream))


def _one(ctx, bvh, host_rays, variant, st):
    d = _rq.DeviceRays.from_host(ctx, host_rays)
    ref = {k: _rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t", "u", "v")}
    out = {k: _rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t", "u", "v")}
    try:
        ctx.trace_config(variant, None)
        ctx.closest_hit_dev(bvh, d, ref)
        for a in out.values():
            a.fill(0x7F)
        ctx.sync()
        used = ctx.bytes_in_use
        _noalloc(ctx, bvh, d, out, st)
        ctx.sync(st)
        assert ctx.bytes_in_use == used, "the no-alloc entry allocated device memory"
        for k, dt in (("prim", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32)):
            g, e = out[k].download(dt, d.n), ref[k].download(dt, d.n)
            assert np.array_equal(g.view(np.uint32), e.view(np.uint32)), k
        assert ctx.device_errors() == 0
        return out["prim"].download(np.int32, d.n)
    finally:
        d.free()
        for a in list(ref.values()) + list(out.values()):
            a.free()


def test_noalloc_equals_closest_hit(ctx, rq, O):
    vs = ctx.trace_variants()
    v12 = [v["index"] for v in vs if v["lds_depth"] == 12][0]
    saved = ctx.trace_config()
    st = ctx.stream_create()
    mesh = random_soup(60000, 27, spread=1.0, size=1.0)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    try:
        big = concat_rays(rays_toward(700000, 91), random_rays(500000, 92))
        ob = O.OracleBVH(mesh, rq.MORTON_SCENE_AABB)
        some = {k: v[::60].copy() for k, v in big.items()}  # 20,000 rays for the oracle
        assert (ob.max_stack(some) > 16).sum() > 2000
        log("no-alloc: 1.2 M rays, 12 LDS entries")
        prim = _one(ctx, b, big, v12, st)
        assert np.array_equal(prim[::60], ob.closest(some)["prim"])
        log("no-alloc: 30,001 rays, default variant")
        small = concat_rays(rays_toward(20001, 93), random_rays(10000, 94))
        _one(ctx, b, small, saved[0], st)
    finally:
        ctx.trace_config(*saved)
        ctx.stream_destroy(st)
        ctx.release(b)
        ctx.release(s)
