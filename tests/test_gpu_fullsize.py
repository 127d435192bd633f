"""Parity at the benchmark workloads' own sizes, against the CPU oracle.

The large-mesh frame (1,310,720 triangles, 1920x1080x4 spp, mirror bounce +
shadow rays):
    every primary ray bit-exact; primary / bounce hit ids and both shadow-flag
    sets identical over the whole frame; sample colours within 2e-5 and the
    image within 1/255; closest hit (prim, t, u, v bit-exact) of every primary
    and bounce ray and any hit of every shadow ray, for every traversal variant.
The raw ray batch (256 L5 icospheres, 5,242,880 triangles, 2^24 uniform
rays): the LBVH
    (codes, order, topology, boxes) bit-exact; every device ray bit-exact;
    closest and any hit on every 16th ray (1,048,576) for every variant, and
    after two treelet rounds.
The 512-mesh batch (36.5 M triangles, 176 M rays, one persistent launch): every
    mesh's LBVH bit-exact; a stride of every mesh's rays (about 2,000 per
    mesh) closest and any hit.
The server (64 resident meshes of the serving distribution, 20.1 M triangles):
    240 Poisson requests (half pinhole, half uniform, 5 % uploads of a new
    mesh into 8 upload slots beside the 64 resident ones) served by
    continuous batching; every request's results on a stride equal the
    oracle's, and the device pinhole generator equals the oracle's.

Semantics checked: main.cpp:343-358 / :373-387 (closest hit, lowest index on
ties), accelerators.h:997-1086 (any hit), main.cpp:451-490 (shading).
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import rq as _rq
from conftest import log
from rq import workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
RAYS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same_rays(got, exp, what):
    for k in RAYS:
        assert np.array_equal(_bits(got[k]), _bits(exp[k])), f"{what}: {k}"


def _same_closest(g, o, what):
    bad = np.nonzero(g["prim"] != o["prim"])[0]
    assert bad.size == 0, f"{what}: {bad.size} hit ids differ, first ray {bad[:5]}"
    for k in ("t", "u", "v"):
        if k in g:
            assert np.array_equal(_bits(g[k]), _bits(o[k])), f"{what}: {k}"


def _take(rays, ids):
    return {k: np.ascontiguousarray(rays[k][ids]) for k in RAYS}


def _export_equal(ctx, bvh, ob, n, what):
    g = ctx.export(bvh, n)
    o = ob.arrays()
    for k in ("codes", "sorted_idx", "left", "right"):
        assert np.array_equal(g[k], o[k]), f"{what}: {k}"
    assert np.array_equal(_bits(g["boxes"]), _bits(o["boxes"])), f"{what}: boxes"


class _keep_trace_config:
    """Restores the context's variants, per-kind refill thresholds and slices."""

    def __init__(self, ctx):
        self.ctx = ctx

    def __enter__(self):
        c = self.ctx
        self.saved = (c.trace_config()[:2], c.trace_refill(), c.trace_slice())
        return self

    def __exit__(self, *a):
        (vc, va), (rc, ra), (sc, sa) = self.saved
        self.ctx.trace_config(vc, va)
        self.ctx.trace_refill(rc, ra)
        self.ctx.trace_slice(sc, sa)


def _variants(any_hit):
    # entry distances on the stack serve the closest-hit pop only
    return [v for v in _rq.Context.trace_variants() if not (any_hit and v["tstack"])]


# ------------------------------------- large-mesh frame (1.31 M triangles)
@pytest.fixture(scope="module")
def c2(ctx, O):
    mesh, p, _ = W.c2()
    log(f"c2: build {mesh.ntri} triangles + one frame {p.width}x{p.height}x{p.spp}")
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, _rq.MORTON_SCENE_AABB)
    r = _rq.Renderer(ctx, p)
    r.frame(b)
    ctx.sync()
    assert ctx.device_errors() == 0
    ob = O.OracleBVH(mesh, _rq.MORTON_SCENE_AABB)
    yield dict(mesh=mesh, p=p, scene=s, bvh=b, ren=r, ob=ob)
    r.close()
    ctx.release(b)
    ctx.release(s)


def test_c2_frame_full_size(ctx, O, c2):
    p, r = c2["p"], c2["ren"]
    n = p.width * p.height * p.spp
    prim = r.rays(0).download()
    orays = O.camera_rays_philox(p.width, p.height, p.spp, p.seed)
    _same_rays(prim, orays, "c2 primary rays")
    log("c2: oracle frame")
    sp = O.shade_params(light=tuple(p.light), emission=tuple(p.emission),
                        surface=tuple(p.surface), sky=tuple(p.sky), bias=p.bias, bounce=p.bounce)
    o = O.render_samples(c2["ob"], c2["mesh"], orays, sp, threads=THREADS)
    g = r.samples()
    for k in ("hit0", "hit1", "shadow0", "shadow1"):
        bad = np.count_nonzero(g[k] != o[k])
        assert bad == 0, f"c2 {k}: {bad} of {n} samples differ"
    cnt = r.counts()
    assert cnt.primary == n
    assert cnt.primary + cnt.bounce + cnt.shadow == int(o["counts"][0])
    assert np.max(np.abs(g["rgb"] - o["rgb"])) < 2e-5
    rgb, ppm = r.image()
    oppm = O.resolve_ppm(o["rgb"], p.width * p.height, p.spp)
    assert np.abs(ppm.astype(np.int16) - oppm.astype(np.int16)).max() <= 1
    # the frame is the workload it is meant to be
    assert (g["hit0"] >= 0).mean() > 0.9 and (g["hit1"] >= 0).sum() > 1000


def test_c2_closest_every_variant(ctx, O, c2):
    """Every primary and bounce ray of the frame, every variant, refill 16 and
    64 (the autotuner's range ends)."""
    b, r = c2["bvh"], c2["ren"]
    batches = {"primary": r.rays(0), "bounce": r.rays(1)}
    ref = {}
    for name, dr in batches.items():
        host = dr.download()
        log(f"c2: oracle closest {name} ({dr.n} rays)")
        ref[name] = c2["ob"].closest(host, threads=THREADS)
    with _keep_trace_config(ctx):
        for v in _variants(False):
            for refill in (16, 64):
                ctx.trace_config(v["index"], None, refill)
                for name, dr in batches.items():
                    log(f"c2: closest {name} variant {v['index']} refill {refill}")
                    hits = {k: _rq.DeviceArray(ctx, 4 * dr.n) for k in ("prim", "t", "u", "v")}
                    ctx.closest_hit_dev(b, dr, hits)
                    ctx.sync()
                    g = {"prim": hits["prim"].download(np.int32, dr.n)}
                    for k in ("t", "u", "v"):
                        g[k] = hits[k].download(np.float32, dr.n)
                    for a in hits.values():
                        a.free()
                    _same_closest(g, ref[name], f"c2 {name} variant {v['index']} refill {refill}")
        assert ctx.device_errors() == 0


def test_c2_any_every_variant(ctx, O, c2):
    b, r = c2["bvh"], c2["ren"]
    dr = r.rays(2)
    host = dr.download()
    log(f"c2: oracle any hit ({dr.n} shadow rays)")
    ref = c2["ob"].any(host, threads=THREADS)["occ"]
    assert ref.sum() > 0
    with _keep_trace_config(ctx):
        for v in _variants(True):
            for refill in (16, 64):
                log(f"c2: any variant {v['index']} refill {refill}")
                ctx.trace_config(None, v["index"], refill)
                occ = _rq.DeviceArray(ctx, dr.n)
                ctx.any_hit_dev(b, dr, occ)
                ctx.sync()
                g = occ.download(np.uint8, dr.n)
                occ.free()
                bad = np.count_nonzero(g != ref)
                assert bad == 0, f"c2 shadow variant {v['index']} refill {refill}: {bad} differ"
        assert ctx.device_errors() == 0


# ------------------------------------ raw ray batch (5.24 M triangles, 2^24 rays)
C3_STRIDE = 16


@pytest.fixture(scope="module")
def c3(ctx, O):
    mesh, (lo, hi), n, seed = W.c3()
    log(f"c3: build {mesh.ntri} triangles, {n} rays")
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, _rq.MORTON_SCENE_AABB)
    rays = ctx.raygen_uniform(n, seed, lo, hi)
    ctx.sync()
    ob = O.OracleBVH(mesh, _rq.MORTON_SCENE_AABB)
    ids = np.arange(0, n, C3_STRIDE, dtype=np.uint32)
    orays = O.uniform_rays_at(ids, seed, lo, hi)
    log(f"c3: oracle closest + any on {ids.size} rays")
    oc = ob.closest(orays, threads=THREADS)
    oa = ob.any(orays, threads=THREADS)["occ"]
    yield dict(mesh=mesh, lo=lo, hi=hi, n=n, seed=seed, scene=s, bvh=b, rays=rays, ob=ob,
               ids=ids, oc=oc, oa=oa)
    rays.free()
    ctx.release(b)
    ctx.release(s)


def test_c3_lbvh_bit_exact(ctx, c3):
    _export_equal(ctx, c3["bvh"], c3["ob"], c3["mesh"].ntri, "c3 lbvh")


def test_c3_device_rays_bit_exact(O, c3):
    g = c3["rays"].download()
    o = O.uniform_rays(c3["n"], c3["seed"], c3["lo"], c3["hi"])
    _same_rays(g, o, "c3 k_raygen_uniform")


def _c3_queries(ctx, c3, variants_closest, variants_any, what):
    b, dr, n, ids = c3["bvh"], c3["rays"], c3["n"], c3["ids"]
    with _keep_trace_config(ctx):
        for v in variants_closest:
            for refill in (16, 64):
                log(f"c3{what}: closest variant {v['index']} refill {refill}")
                ctx.trace_config(v["index"], None, refill)
                hits = {k: _rq.DeviceArray(ctx, 4 * n) for k in ("prim", "t", "u", "v")}
                ctx.closest_hit_dev(b, dr, hits)
                ctx.sync()
                g = {"prim": hits["prim"].download(np.int32, n)[ids]}
                for k in ("t", "u", "v"):
                    g[k] = hits[k].download(np.float32, n)[ids]
                for a in hits.values():
                    a.free()
                _same_closest(g, c3["oc"], f"c3{what} closest variant {v['index']} refill {refill}")
        for v in variants_any:
            for refill in (16, 64):
                log(f"c3{what}: any variant {v['index']} refill {refill}")
                ctx.trace_config(None, v["index"], refill)
                occ = _rq.DeviceArray(ctx, n)
                ctx.any_hit_dev(b, dr, occ)
                ctx.sync()
                g = occ.download(np.uint8, n)[ids]
                occ.free()
                bad = np.count_nonzero(g != c3["oa"])
                assert bad == 0, f"c3{what} any variant {v['index']} refill {refill}: {bad} differ"
        assert ctx.device_errors() == 0


def test_c3_closest_and_any_every_variant(ctx, c3):
    assert (c3["oc"]["prim"] >= 0).sum() > 10000 and c3["oa"].sum() > 10000
    _c3_queries(ctx, c3, _variants(False), _variants(True), "")


def test_c3_after_treelets(ctx, c3):
    """Two treelet rounds restructure the tree in place; the answers (which
    do not depend on the topology) stay the oracle's."""
    log("c3: two treelet rounds")
    ctx.treelets(c3["bvh"], 2)
    ctx.sync()
    wide = [v for v in _rq.Context.trace_variants() if v["wide"]]
    _c3_queries(ctx, c3, [wide[0], wide[-1]], [wide[0], wide[-1]], " treelets")


# ----------------------------------------------------- 512-mesh batch
C4_RAYS_PER_MESH = 2048


def test_c4_full_workload(ctx, O):
    tris, nrays = W.c4_sizes()
    count = len(tris)
    log(f"c4: generate {count} meshes")
    meshes = [W.c4_mesh(m, tris[m]) for m in range(count)]
    boxes = [W.inflated_box(m) for m in meshes]
    items, scenes = [], []
    for m, mesh in enumerate(meshes):
        n = int(nrays[m])
        s = ctx.upload_scene(mesh)
        items.append(dict(scene=s, rays=ctx.raygen_uniform(n, W.c4_ray_seed(m), *boxes[m]),
                          hits={k: _rq.DeviceArray(ctx, 4 * n) for k in ("prim", "t", "u", "v")},
                          occ=_rq.DeviceArray(ctx, n)))
        scenes.append(s)
    log(f"c4: {sum(x.ntri for x in meshes)} triangles, {int(nrays.sum())} rays; batch run")
    batch = _rq.Batch(ctx, items, _rq.MORTON_SCENE_AABB)
    try:
        batch.run_graph()
        ctx.sync()
        assert batch.errors() == 0
        log("c4: oracle builds")
        with cf.ThreadPoolExecutor(THREADS) as ex:
            obs = list(ex.map(lambda x: O.OracleBVH(x, _rq.MORTON_SCENE_AABB), meshes))
        for m in range(count):
            if m % 64 == 0:
                log(f"c4: trees {m}..")
            _export_equal(ctx, batch.bvh(m), obs[m], meshes[m].ntri, f"c4 mesh {m}")
        sel = {}
        for m in range(count):
            n = int(nrays[m])
            sel[m] = np.unique(np.linspace(0, n - 1, C4_RAYS_PER_MESH).astype(np.uint32))
        log("c4: oracle queries")

        def oracle_q(m):
            r = O.uniform_rays_at(sel[m], W.c4_ray_seed(m), *boxes[m])
            return r, obs[m].closest(r, threads=1), obs[m].any(r, threads=1)["occ"]

        with cf.ThreadPoolExecutor(THREADS) as ex:
            oq = list(ex.map(oracle_q, range(count)))
        hits_total = 0
        for m in range(count):
            n, it, ids = int(nrays[m]), items[m], sel[m]
            g = {"prim": it["hits"]["prim"].download(np.int32, n)[ids]}
            for k in ("t", "u", "v"):
                g[k] = it["hits"][k].download(np.float32, n)[ids]
            _same_closest(g, oq[m][1], f"c4 mesh {m}")
            occ = it["occ"].download(np.uint8, n)[ids]
            assert np.array_equal(occ, oq[m][2]), f"c4 mesh {m} any hit"
            hits_total += int((g["prim"] >= 0).sum())
            if m % 97 == 0:  # device rays of a few meshes, every one bit-exact
                dr = it["rays"].download()
                o = O.uniform_rays(n, W.c4_ray_seed(m), *boxes[m])
                _same_rays(dr, o, f"c4 mesh {m} rays")
        assert hits_total > 10000
    finally:
        batch.close()
        for it in items:
            it["rays"].free()
            it["occ"].free()
            for a in it["hits"].values():
                a.free()
        for s in scenes:
            ctx.release(s)


# -------------------------------------------------- server, 64 meshes
C5_REQUESTS = 240
C5_RAYS_PER_REQUEST = 4096
C5_UPLOAD_SLOTS = 8


def test_c5_served_requests_equal_the_oracle(O):
    resident = W.c5_meshes(64, 5)
    uploads = W.c5_meshes(8, 6)
    bounds = {id(m): m.bounds() for m in resident + uploads}
    rng = np.random.default_rng(17)
    arrival = np.cumsum(rng.exponential(1 / 2000.0, C5_REQUESTS))
    sizes = np.rint(np.exp(rng.uniform(np.log(4096.0), np.log(4194304.0), C5_REQUESTS)))
    sizes = sizes.astype(np.int64)
    target = rng.integers(0, len(resident), C5_REQUESTS)
    upload = rng.random(C5_REQUESTS) < 0.05
    upload[[7, 100]] = True            # at least two, whatever the draw
    log(f"c5: {sum(m.ntri for m in resident)} resident triangles, {C5_REQUESTS} requests, "
        f"{int(sizes.sum())} rays, {int(upload.sum())} uploads")
    # 64 resident slots plus 8 that the uploads recycle
    srv = _rq.Server(0, mesh_slots=len(resident) + C5_UPLOAD_SLOTS)
    try:
        slots = [srv.load_mesh(m) for m in resident]
        specs, outs, tickets = [], [], []
        t0 = _rq.server_clock() + 0.02
        for j in range(C5_REQUESTS):
            mesh = uploads[j % len(uploads)] if upload[j] else resident[target[j]]
            sp = W.c5_rays_spec(bounds[id(mesh)], j % 2, 500 + j)
            q = _rq.Q_CLOSEST | (_rq.Q_ANY if j % 3 == 0 else 0)
            n = int(sizes[j])
            o = dict(prim=np.empty(n, np.int32), t=np.empty(n, np.float32),
                     occ=np.empty(n, np.uint8) if q & _rq.Q_ANY else None)
            while _rq.server_clock() < t0 + arrival[j]:
                pass
            kw = dict(new_mesh=mesh) if upload[j] else dict(mesh=slots[target[j]])
            kw.update({k: sp[k] for k in ("pinhole", "box", "seed") if k in sp})
            tickets.append(srv.submit(n, q, prim=o["prim"], t=o["t"], occ=o["occ"], **kw))
            specs.append((mesh, sp, q, n))
            outs.append(o)
        log("c5: submitted; waiting")
        st = [srv.wait(tk) for tk in tickets]
        assert all(s.state == _rq.REQ_DONE for s in st)
        stats = srv.stats()
        assert stats["failed"] == 0 and stats["max_step_requests"] >= 2
        # a few requests again, alone: identical results
        for j in (0, 7, 33, 100):
            mesh, sp, q, n = specs[j]
            o = dict(prim=np.empty(n, np.int32), t=np.empty(n, np.float32),
                     occ=np.empty(n, np.uint8) if q & _rq.Q_ANY else None)
            kw = dict(new_mesh=mesh) if upload[j] else dict(mesh=slots[target[j]])
            kw.update({k: sp[k] for k in ("pinhole", "box", "seed") if k in sp})
            log(f"c5: request {j} alone")
            srv.wait(srv.submit(n, q, prim=o["prim"], t=o["t"], occ=o["occ"], **kw))
            assert np.array_equal(o["prim"], outs[j]["prim"]), j
            assert np.array_equal(_bits(o["t"]), _bits(outs[j]["t"])), j
            if o["occ"] is not None:
                assert np.array_equal(o["occ"], outs[j]["occ"]), j
    finally:
        srv.close()
    log("c5: oracle")
    with cf.ThreadPoolExecutor(THREADS) as ex:
        obs = dict(zip([id(m) for m in resident + uploads],
                       ex.map(lambda x: O.OracleBVH(x, _rq.MORTON_SCENE_AABB),
                              resident + uploads)))

    def check(j):
        mesh, sp, q, n = specs[j]
        ids = np.unique(np.linspace(0, n - 1, C5_RAYS_PER_REQUEST).astype(np.uint32))
        if "pinhole" in sp:
            c = sp["pinhole"]
            r = O.pinhole_rays_at(ids, n, sp["seed"], c["eye"], c["look_at"], c["up"],
                                  c["fov_deg"], c["width"])
        else:
            r = O.uniform_rays_at(ids, sp["seed"], *sp["box"])
        oc = obs[id(mesh)].closest(r, threads=1)
        g = dict(prim=outs[j]["prim"][ids], t=outs[j]["t"][ids])
        _same_closest(g, oc, f"c5 request {j}")
        if q & _rq.Q_ANY:
            oa = obs[id(mesh)].any(r, threads=1)["occ"]
            assert np.array_equal(outs[j]["occ"][ids], oa), f"c5 request {j} any hit"
        return int((oc["prim"] >= 0).sum())

    with cf.ThreadPoolExecutor(THREADS) as ex:
        hits = list(ex.map(check, range(C5_REQUESTS)))
    assert sum(hits) > 10000


def test_device_pinhole_generator_equals_the_oracle(ctx, O):
    """k_raygen_pinhole (the serving requests' camera) against the oracle's
    restatement, every ray, ragged last row and a tilted basis."""
    cases = [(100003, 9, (1.0, 2.0, 3.0), (1.5, 1.0, -4.0), (0.0, 1.0, 0.0), 40.0, 1024),
             (4096, 1, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.3, 1.0, 0.1), 30.0, 64),
             (1, 2, (5.0, -1.0, 2.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 75.0, 1)]
    for n, seed, eye, look, up, fov, w in cases:
        log(f"pinhole generator n={n}")
        dr = ctx.raygen_pinhole(n, seed, eye, look, up, fov, w)
        ctx.sync()
        g = dr.download()
        dr.free()
        o = O.pinhole_rays_at(None, n, seed, eye, look, up, fov, w)
        _same_rays(g, o, f"pinhole n={n}")
