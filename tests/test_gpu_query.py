"""GPU closest-hit / any-hit parity with the CPU oracle: identical hit ids
(lowest index on ties), t/u/v bit-exact, identical traversal counts, and
equality with brute force (the reference's NONE loop, main.cpp:373-387)."""
import numpy as np
import pytest

from conftest import log
from helpers import (concat_rays, duplicate_mesh, edge_rays, mesh_configs, nonfinite_rays,
                     random_rays, random_soup, rays_toward)

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _query_parity(ctx, O, mesh, rays, mode=1, brute=True):
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, mode)
    try:
        ob = O.OracleBVH(mesh, mode)
        g = ctx.closest_hit(b, rays, counted=True)
        o = ob.closest(rays)
        assert np.array_equal(g["prim"], o["prim"]), "closest prim ids"
        assert np.array_equal(_bits(g["t"]), _bits(o["t"])), "closest t"
        assert np.array_equal(_bits(g["u"]), _bits(o["u"])), "closest u"
        assert np.array_equal(_bits(g["v"]), _bits(o["v"])), "closest v"
        assert np.array_equal(g["visits"], o["visits"]), "node visits"
        assert np.array_equal(g["tests"], o["tests"]), "triangle tests"
        g2 = ctx.closest_hit(b, rays)  # uncounted kernel, same answers
        assert np.array_equal(g2["prim"], o["prim"])
        assert np.array_equal(_bits(g2["t"]), _bits(o["t"]))
        ga = ctx.any_hit(b, rays, counted=True)
        oa = ob.any(rays)
        assert np.array_equal(ga["occ"], oa["occ"]), "any-hit"
        assert np.array_equal(ga["visits"], oa["visits"])
        assert np.array_equal(ga["tests"], oa["tests"])
        ga2 = ctx.any_hit(b, rays)
        assert np.array_equal(ga2["occ"], oa["occ"])
        if brute:
            bf = O.closest_brute(mesh, rays)
            assert np.array_equal(o["prim"], bf["prim"]), "oracle BVH vs brute force"
            assert np.array_equal(_bits(o["t"]), _bits(bf["t"]))
            ab = O.any_brute(mesh, rays)
            assert np.array_equal(oa["occ"], ab)
        return g
    finally:
        ctx.release(b)
        ctx.release(s)


@pytest.mark.parametrize("name", list(mesh_configs()))
def test_query_small(ctx, O, name):
    mesh = mesh_configs()[name]()
    rays = concat_rays(rays_toward(3000, 11), random_rays(2000, 12), edge_rays())
    log(f"query {name} n={len(rays['ox'])}")
    g = _query_parity(ctx, O, mesh, rays)
    assert (g["prim"] >= 0).any()


def test_query_ties_lowest_index(ctx, O):
    """Identical triangles: every hit is a tie, the lowest index must win."""
    mesh = duplicate_mesh(300)
    rays = rays_toward(2000, 5, spread=0.5)
    log("query ties")
    g = _query_parity(ctx, O, mesh, rays)
    hit = g["prim"] >= 0
    assert hit.sum() > 1000
    assert (g["prim"][hit] == 0).all()


def test_query_ref_affine_mode(ctx, O, rq):
    """REF_AFFINE Morton map leaves few distinct codes (accelerators.h:577):
    the tree degenerates towards index order, results must still be exact."""
    mesh = rq.icosphere(4, 3, center=(0, 0, -5))
    rays = rays_toward(4000, 21)
    log("query ref-affine")
    _query_parity(ctx, O, mesh, rays, mode=0)


def test_query_zero_rays(ctx, O, rq):
    mesh = rq.icosphere(1, 1, center=(0, 0, -5))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    rays = {k: np.zeros(0, np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin",
                                                    "tmax")}
    log("query zero rays")
    g = ctx.closest_hit(b, rays)
    assert g["prim"].size == 0
    ctx.release(b)
    ctx.release(s)


def test_query_empty_scene_misses(ctx, rq):
    m = rq.MeshArrays(np.zeros(0, np.float32), np.zeros(0, np.float32),
                      np.zeros(0, np.float32), np.zeros(0, np.uint32))
    s = ctx.upload_scene(m)
    b = ctx.build_lbvh(s, 1)
    log("query empty scene")
    g = ctx.closest_hit(b, rays_toward(500, 1))
    assert (g["prim"] == -1).all() and np.isinf(g["t"]).all()
    a = ctx.any_hit(b, rays_toward(500, 1))
    assert (a["occ"] == 0).all()
    ctx.release(b)
    ctx.release(s)


def test_query_c1_camera(ctx, O, rq):
    """The small frame's mesh (81,920 triangles) with the reference camera rays at 640x480."""
    mesh = rq.icosphere(6, 1, center=(0, 0, -5))
    jit = O.mt19937_jitter(2 * 640 * 480)
    rays = O.camera_rays_ref(640, 480, 1, jit)
    log("query c1 camera")
    g = _query_parity(ctx, O, mesh, rays, brute=False)
    assert (g["prim"] >= 0).mean() > 0.2
    # brute force on a strided sample
    sub = {k: v[::97].copy() for k, v in rays.items()}
    bf = O.closest_brute(mesh, sub)
    assert np.array_equal(g["prim"][::97], bf["prim"])


def test_query_soup_random(ctx, O):
    mesh = random_soup(20000, 9, spread=1.0, size=0.03)
    rays = random_rays(20000, 4, lo=(-1.2, -1.2, -1.2), hi=(1.2, 1.2, 1.2))
    log("query soup random")
    _query_parity(ctx, O, mesh, rays, brute=False)
    sub = {k: v[::50].copy() for k, v in rays.items()}
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    g = ctx.closest_hit(b, sub)
    bf = O.closest_brute(mesh, sub)
    assert np.array_equal(g["prim"], bf["prim"])
    ctx.release(b)
    ctx.release(s)


def test_all_trace_variants_identical(ctx, O, rq):
    """Every traversal kernel variant (block size, LDS stack depth, entry
    distances on the stack, refill threshold) returns the oracle's answers."""
    mesh = rq.icosphere(5, 4, center=(0, 0, -5))
    rays = concat_rays(rays_toward(6000, 31), random_rays(3000, 32), edge_rays())
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    ob = O.OracleBVH(mesh, 1)
    o = ob.closest(rays)
    oa = ob.any(rays)
    saved = ctx.trace_config()
    try:
        for v in ctx.trace_variants():
            for refill in (1, 16, 64):
                log(f"variant {v} refill {refill}")
                ctx.trace_config(v["index"], v["index"] if not v["tstack"] else 0, refill)
                g = ctx.closest_hit(b, rays)
                assert np.array_equal(g["prim"], o["prim"]), v
                assert np.array_equal(_bits(g["t"]), _bits(o["t"])), v
                assert np.array_equal(_bits(g["u"]), _bits(o["u"])), v
                ga = ctx.any_hit(b, rays)
                assert np.array_equal(ga["occ"], oa["occ"]), v
        assert ctx.device_errors() == 0
    finally:
        ctx.trace_config(*saved)
        ctx.release(b)
        ctx.release(s)


def test_nonfinite_rays_every_variant(ctx, O, rq):
    """Rays with NaN or infinite components miss in the oracle and in brute
    force (no finite t exists), and every kernel variant, binary and 4-wide,
    returns the oracle's answers for them and for the finite rays beside them
    (a ray whose every axis is NaN passes every box, the empty slots of the
    wide nodes included, which their zero links keep out of the walk)."""
    mesh = rq.icosphere(3, 5, center=(0, 0, -5))
    rays = concat_rays(nonfinite_rays(), rays_toward(500, 41))
    finite = np.isfinite(np.stack([rays[k] for k in ("ox", "oy", "oz", "dx", "dy", "dz")])).all(0)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    ob = O.OracleBVH(mesh, 1)
    o = ob.closest(rays)
    oa = ob.any(rays)
    bf = O.closest_brute(mesh, rays)
    assert (o["prim"][~finite] == -1).all() and not oa["occ"][~finite].any()
    assert np.array_equal(o["prim"], bf["prim"])
    assert (o["prim"][finite] >= 0).any()
    saved = ctx.trace_config()
    try:
        for v in ctx.trace_variants():
            log(f"non-finite rays, variant {v}")
            ctx.trace_config(v["index"], v["index"] if not v["tstack"] else 0, 16)
            g = ctx.closest_hit(b, rays)
            assert np.array_equal(g["prim"], o["prim"]), v
            assert np.array_equal(_bits(g["t"][finite]), _bits(o["t"][finite])), v
            ga = ctx.any_hit(b, rays)
            assert np.array_equal(ga["occ"], oa["occ"]), v
        assert ctx.device_errors() == 0
    finally:
        ctx.trace_config(*saved)
        ctx.release(b)
        ctx.release(s)


def test_ray_stream_slices_and_counter_reuse(ctx, O, rq):
    """Every ray of a launch is traced exactly once whatever the slice size
    (large slices, then quarter ones, from one counter per launch slot that is
    never reset): launches large enough that the waves take slices from the
    counter, small ones covered by the waves' first slices, and more launches
    than there are counter slots, each against the oracle."""
    mesh = rq.icosphere(4, 1, center=(0, 0, -5))
    n_big = 2_500_000
    rays = rays_toward(n_big, 51)
    ob = O.OracleBVH(mesh, 1)
    o, oa = ob.closest(rays), ob.any(rays)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    drays = rq.DeviceRays.from_host(ctx, rays)
    hits = {k: rq.DeviceArray(ctx, 4 * n_big) for k in ("prim", "t")}
    occ = rq.DeviceArray(ctx, n_big)
    saved, saved_slice = ctx.trace_config(), ctx.trace_slice()

    def check(n, any_hit):
        drays.n = n
        if any_hit:
            occ.fill(0x7F)
            ctx.any_hit_dev(b, drays, occ)
            assert np.array_equal(occ.download(np.uint8, n), oa["occ"][:n]), n
        else:
            hits["prim"].fill(0x7F)
            ctx.closest_hit_dev(b, drays, hits)
            assert np.array_equal(hits["prim"].download(np.int32, n), o["prim"][:n]), n
            assert np.array_equal(_bits(hits["t"].download(np.float32, n)), _bits(o["t"][:n])), n

    try:
        for sl in (64, 256, 4096):
            log(f"ray stream slice {sl}")
            ctx.trace_slice(sl, sl)
            for n in (n_big, 300_000, 1, 100_000, n_big - 12345):
                check(n, False)
                check(n, True)
        ctx.trace_slice(1024, 1024)
        log("ray stream counter reuse")
        for k in range(80):
            check(300_000 + 4099 * k, k % 2 == 1)
        assert ctx.device_errors() == 0
    finally:
        drays.n = n_big
        ctx.trace_config(*saved)
        ctx.trace_slice(*saved_slice)
        drays.free()
        for a in hits.values():
            a.free()
        occ.free()
        ctx.release(b)
        ctx.release(s)


def test_fresh_context_defaults(rq):
    """A new context walks the 4-wide nodes (variant 9 closest, 14 any hit),
    refills once every lane is idle and takes 1024-ray slices."""
    c = rq.Context(0)
    try:
        assert c.trace_config() == (9, 14, 64)
        assert c.trace_refill() == (64, 64)
        assert c.trace_slice() == (1024, 1024)
        v = c.trace_variants()
        assert v[9]["wide"] and v[14]["wide"] and not v[14]["tstack"]
    finally:
        c.close()


def test_trace_slice_refused(ctx, rq):
    for bad in ((0, 1024), (32, 1024), (1024, 8192), (1000, 1024)):
        with pytest.raises(rq.RQError):
            ctx.trace_slice(*bad)


def test_autotune_picks_a_valid_variant(ctx, rq):
    mesh = rq.icosphere(5, 4, center=(0, 0, -5))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    saved = ctx.trace_config()
    rays = rq.DeviceRays.from_host(ctx, rays_toward(100000, 3))
    try:
        log("autotune")
        v, refill, ms = ctx.autotune(b, rays, any_hit=False, repeats=2)
        assert 0 <= v < len(ctx.trace_variants()) and refill in (16, 48, 56, 64) and ms > 0
        assert ctx.trace_config()[0] == v
        va, ra, _ = ctx.autotune(b, rays, any_hit=True, repeats=2)
        assert not ctx.trace_variants()[va]["tstack"]
        # each query kind keeps its own tuned refill threshold
        assert ctx.trace_refill() == (refill, ra)
        assert ctx.trace_config()[:2] == (v, va)
        assert all(x in (256, 512, 1024, 2048) for x in ctx.trace_slice())
    finally:
        rays.free()
        ctx.trace_config(*saved)
        ctx.trace_slice(1024, 1024)
        ctx.release(b)
        ctx.release(s)


def _hip():
    """The HIP runtime, for stream and graph calls around the library's own."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    vp = C.c_void_p
    for name, args in (("hipStreamCreate", [C.POINTER(vp)]), ("hipStreamDestroy", [vp]),
                       ("hipStreamBeginCapture", [vp, C.c_int]),
                       ("hipStreamEndCapture", [vp, C.POINTER(vp)]),
                       ("hipGraphInstantiate", [C.POINTER(vp), vp, vp, vp, C.c_size_t]),
                       ("hipGraphLaunch", [vp, vp]), ("hipGraphExecDestroy", [vp]),
                       ("hipGraphDestroy", [vp])):
        getattr(hip, name).argtypes = args
        getattr(hip, name).restype = C.c_int
    return hip


def test_concurrent_streams_deep_stacks(ctx, O, rq):
    """Queries on two streams at once over a scene of large overlapping
    triangles, in binary variants with 12 (entry distances kept) and 16 LDS
    stack entries: the oracle's walk, which pushes exactly where the binary
    kernel does, needs more than 16 entries on tens of thousands of these
    rays, so the launches spill to the global overflow area; each stream has
    an area of its own, and every answer equals the oracle's."""
    mesh = random_soup(60000, 27, spread=1.0, size=1.0)
    ra = concat_rays(rays_toward(30000, 81), random_rays(20000, 82))
    rb = concat_rays(rays_toward(30000, 83), random_rays(20000, 84))
    ob = O.OracleBVH(mesh, 1)
    oa, obb = ob.closest(ra), ob.closest(rb)
    deep = [(ob.max_stack(r) > 16).sum() for r in (ra, rb)]
    assert min(deep) > 1000, deep
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    da, db = rq.DeviceRays.from_host(ctx, ra), rq.DeviceRays.from_host(ctx, rb)
    ha = {k: rq.DeviceArray(ctx, 4 * da.n) for k in ("prim", "t")}
    hb = {k: rq.DeviceArray(ctx, 4 * db.n) for k in ("prim", "t")}
    st = [ctx.stream_create(), ctx.stream_create()]
    saved = ctx.trace_config()
    vs = ctx.trace_variants()
    v12 = [v["index"] for v in vs if v["lds_depth"] == 12][0]
    v16 = [v["index"] for v in vs if v["lds_depth"] == 16 and not v["tstack"] and not v["wide"]][0]
    try:
        for var in (v12, v16):
            ctx.trace_config(var, 0, 64)
            for rep in range(3):
                log(f"deep stacks: variant {var} rep {rep}")
                for h in (ha, hb):
                    h["prim"].fill(0x7F)
                ctx.closest_hit_dev(b, da, ha, stream=st[0])
                ctx.closest_hit_dev(b, db, hb, stream=st[1])
                ctx.sync(st[0])
                ctx.sync(st[1])
                for h, o, n in ((ha, oa, da.n), (hb, obb, db.n)):
                    assert np.array_equal(h["prim"].download(np.int32, n), o["prim"]), rep
                    assert np.array_equal(_bits(h["t"].download(np.float32, n)), _bits(o["t"]))
        assert ctx.device_errors() == 0
    finally:
        ctx.trace_config(*saved)
        for x in st:
            ctx.stream_destroy(x)
        for x in (da, db):
            x.free()
        for h in (ha, hb):
            for a in h.values():
                a.free()
        ctx.release(b)
        ctx.release(s)


def test_seventeenth_caller_stream(ctx, rq):
    """Caller-owned streams beyond the per-context limit of areas (16): a
    query on a new stream takes over the area of a stream whose work has
    completed instead of failing; rq_stream_forget frees an area."""
    import ctypes as C

    hip = _hip()
    mesh = random_soup(500, 29, spread=1.5)
    rays = rays_toward(4000, 86)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    d = rq.DeviceRays.from_host(ctx, rays)
    h = {k: rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t")}
    ctx.closest_hit_dev(b, d, h)
    ctx.sync()
    ref = h["prim"].download(np.int32, d.n)
    streams = []
    try:
        for k in range(24):
            x = C.c_void_p()
            assert hip.hipStreamCreate(C.byref(x)) == 0
            streams.append(x)
            h["prim"].fill(0x7F)
            ctx.closest_hit_dev(b, d, h, stream=x.value)
            ctx.sync(x.value)
            assert np.array_equal(h["prim"].download(np.int32, d.n), ref), k
        assert ctx.device_errors() == 0
    finally:
        ctx.sync()
        for x in streams:
            ctx.stream_forget(x.value)
            hip.hipStreamDestroy(x)
        d.free()
        for a in h.values():
            a.free()
        ctx.release(b)
        ctx.release(s)


def test_captured_query_replays(ctx, O, rq):
    """Closest-hit and any-hit queries captured into a HIP graph on a caller
    stream, replayed four times with eager queries on the same stream and
    synchronisations between the replays: every replay equals the oracle
    (a captured query zeroes its own ray-stream counter inside the graph)."""
    import ctypes as C

    hip = _hip()
    mesh = rq.icosphere(6, 5, center=(0, 0, -5))
    rays = concat_rays(rays_toward(300000, 87), random_rays(200000, 88))
    ob = O.OracleBVH(mesh, 1)
    oc, oa = ob.closest(rays), ob.any(rays)["occ"]
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    d = rq.DeviceRays.from_host(ctx, rays)
    h = {k: rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t")}
    occ = rq.DeviceArray(ctx, d.n)
    he = {k: rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t")}
    st = ctx.stream_create()
    graph, exe = C.c_void_p(), C.c_void_p()
    try:
        ctx.closest_hit_dev(b, d, he, stream=st)  # the stream's overflow area exists
        ctx.sync(st)
        assert hip.hipStreamBeginCapture(C.c_void_p(st), 0) == 0  # hipStreamCaptureModeGlobal
        ctx.closest_hit_dev(b, d, h, stream=st)
        ctx.any_hit_dev(b, d, occ, stream=st)
        assert hip.hipStreamEndCapture(C.c_void_p(st), C.byref(graph)) == 0
        assert hip.hipGraphInstantiate(C.byref(exe), graph, None, None, 0) == 0
        for rep in range(4):
            log(f"captured query replay {rep}")
            h["prim"].fill(0x7F)
            occ.fill(0x7F)
            ctx.sync()
            assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
            ctx.closest_hit_dev(b, d, he, stream=st)  # eager work between the replays
            ctx.sync(st)
            assert np.array_equal(h["prim"].download(np.int32, d.n), oc["prim"]), rep
            assert np.array_equal(_bits(h["t"].download(np.float32, d.n)), _bits(oc["t"])), rep
            assert np.array_equal(occ.download(np.uint8, d.n), oa), rep
            assert np.array_equal(he["prim"].download(np.int32, d.n), oc["prim"]), rep
        assert ctx.device_errors() == 0
    finally:
        if exe.value:
            hip.hipGraphExecDestroy(exe)
        if graph.value:
            hip.hipGraphDestroy(graph)
        ctx.stream_destroy(st)
        d.free()
        occ.free()
        for a in list(h.values()) + list(he.values()):
            a.free()
        ctx.release(b)
        ctx.release(s)


def test_capture_on_a_new_stream_is_refused(ctx, rq):
    """The first query on a stream allocates its overflow area, which is not
    allowed under capture: RQ_E_UNSUPPORTED, and the capture can be ended."""
    import ctypes as C

    hip = _hip()
    mesh = random_soup(300, 30, spread=1.5)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    d = rq.DeviceRays.from_host(ctx, rays_toward(1000, 89))
    h = {k: rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t")}
    st = ctx.stream_create()
    graph = C.c_void_p()
    try:
        assert hip.hipStreamBeginCapture(C.c_void_p(st), 2) == 0  # relaxed mode
        with pytest.raises(rq.RQError) as e:
            ctx.closest_hit_dev(b, d, h, stream=st)
        assert e.value.code == rq.RQ_E_UNSUPPORTED
        hip.hipStreamEndCapture(C.c_void_p(st), C.byref(graph))
        if graph.value:
            hip.hipGraphDestroy(graph)
        ctx.closest_hit_dev(b, d, h, stream=st)  # eager: fine
        ctx.sync(st)
        assert (h["prim"].download(np.int32, d.n) >= 0).any()
    finally:
        ctx.stream_destroy(st)
        d.free()
        for a in h.values():
            a.free()
        ctx.release(b)
        ctx.release(s)


def test_stream_overflow_areas_are_released(ctx, rq):
    """A stream's stack overflow area is freed with the stream: many streams
    created, used and destroyed in turn never reach the per-context limit."""
    mesh = random_soup(500, 29, spread=1.5)
    rays = rays_toward(4000, 85)
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, 1)
    d = rq.DeviceRays.from_host(ctx, rays)
    h = {k: rq.DeviceArray(ctx, 4 * d.n) for k in ("prim", "t")}
    used = ctx.bytes_in_use
    try:
        for _ in range(40):
            x = ctx.stream_create()
            ctx.closest_hit_dev(b, d, h, stream=x)
            ctx.stream_destroy(x)
        assert ctx.bytes_in_use == used
    finally:
        d.free()
        for a in h.values():
            a.free()
        ctx.release(b)
        ctx.release(s)
