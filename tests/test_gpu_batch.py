"""Batched build + traversal (one persistent launch over many meshes) against
the per-mesh path: for every mesh the batch's LBVH (codes, order, topology,
boxes) and its closest-hit / any-hit answers are bit-identical to
rq_build_lbvh + rq_closest_hit / rq_any_hit on that mesh alone, which the
query and build tests pin to the CPU oracle; two meshes are also checked
against the oracle directly."""
import numpy as np
import pytest

import rq as _rq
from conftest import log
from helpers import concat_rays, duplicate_mesh, edge_rays, random_rays, random_soup, rays_toward

pytestmark = pytest.mark.gpu

# (name, mesh builder, rays: (toward, random) counts, closest?, any?)
CASES = [
    ("tri1", lambda: random_soup(1, 5, spread=0.2, size=1.0), (300, 200), True, True),
    ("tri2", lambda: random_soup(2, 6, spread=0.3, size=0.8), (0, 0), True, True),
    ("dup64", lambda: duplicate_mesh(64), (1500, 500), True, True),
    ("soup100", lambda: random_soup(100, 1, spread=1.5), (2000, 1000), True, False),
    ("ico2", lambda: _rq.icosphere(2, 1, center=(0, 0, -5)), (1, 0), True, True),
    ("soup2048", lambda: random_soup(2048, 7, spread=1.5, size=0.03), (3000, 3000), False, True),
    ("soup2049", lambda: random_soup(2049, 8, spread=1.5, size=0.03), (4000, 2000), True, True),
    ("soup4097", lambda: random_soup(4097, 3, spread=1.5, size=0.02), (9000, 5000), True, True),
    ("ico5", lambda: _rq.icosphere(5, 2, center=(0, 0, -5)), (20000, 13333), True, True),
    ("soup70k", lambda: random_soup(70001, 9, spread=1.8, size=0.01), (30000, 30000), True, True),
    ("ico4", lambda: _rq.icosphere(4, 3, center=(0.3, -0.2, -5.5)), (5000, 0), True, True),
]
ORACLE_CASES = ("soup100", "ico2", "soup4097")


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


class Setup:
    def __init__(self, ctx, mode=_rq.MORTON_SCENE_AABB, cases=CASES):
        self.ctx = ctx
        self.mode = mode
        self.meshes, self.rays, self.items, self.scenes, self.names = [], [], [], [], []
        for k, (name, make, (na, nb), closest, anyq) in enumerate(cases):
            mesh = make()
            parts = []
            if na:
                parts.append(rays_toward(na, 100 + k))
            if nb:
                parts.append(random_rays(nb, 200 + k))
            if na and k % 3 == 0:
                parts.append(edge_rays())
            rays = concat_rays(*parts) if parts else None
            s = ctx.upload_scene(mesh)
            n = len(rays["ox"]) if rays is not None else 0
            drays = _rq.DeviceRays.from_host(ctx, rays) if n else None
            hits = ({f: _rq.DeviceArray(ctx, 4 * max(n, 1)) for f in ("prim", "t", "u", "v")}
                    if closest else None)
            occ = _rq.DeviceArray(ctx, max(n, 1)) if anyq else None
            self.meshes.append(mesh)
            self.rays.append(rays)
            self.scenes.append(s)
            self.names.append(name)
            self.items.append(dict(scene=s, rays=drays, hits=hits, occ=occ))
        self.batch = _rq.Batch(ctx, self.items, mode)
        self.ref = None

    def clear(self):
        for it in self.items:
            if it["hits"]:
                for a in it["hits"].values():
                    a.fill(0xA5)
            if it["occ"]:
                it["occ"].fill(0x7F)
        self.ctx.sync()

    def results(self):
        self.ctx.sync()
        out = []
        for it, rays in zip(self.items, self.rays):
            n = it["rays"].n if it["rays"] is not None else 0
            r = {}
            if it["hits"]:
                r["prim"] = it["hits"]["prim"].download(np.int32, n)
                for f in ("t", "u", "v"):
                    r[f] = it["hits"][f].download(np.float32, n)
            if it["occ"]:
                r["occ"] = it["occ"].download(np.uint8, n)
            out.append(r)
        return out

    def per_mesh(self):
        """Per-mesh path on the same scenes: build, query, export."""
        if self.ref is not None:
            return self.ref
        ref = []
        for s, mesh, rays, it in zip(self.scenes, self.meshes, self.rays, self.items):
            b = self.ctx.build_lbvh(s, self.mode)
            r = {"export": self.ctx.export(b, mesh.ntri), "info": self.ctx.bvh_info(b)}
            if rays is not None:
                if it["hits"]:
                    r.update(self.ctx.closest_hit(b, rays))
                if it["occ"]:
                    r["occ"] = self.ctx.any_hit(b, rays)["occ"]
            self.ctx.release(b)
            ref.append(r)
        self.ref = ref
        return ref

    def check(self, got, what):
        ref = self.per_mesh()
        for i, (g, r, name) in enumerate(zip(got, ref, self.names)):
            if self.rays[i] is None:
                continue
            if "prim" in g:
                assert np.array_equal(g["prim"], r["prim"]), f"{what} {name}: closest prim"
                for f in ("t", "u", "v"):
                    assert np.array_equal(_bits(g[f]), _bits(r[f])), f"{what} {name}: {f}"
            if "occ" in g:
                assert np.array_equal(g["occ"], r["occ"]), f"{what} {name}: any-hit"

    def check_trees(self):
        ref = self.per_mesh()
        for i, (r, mesh, name) in enumerate(zip(ref, self.meshes, self.names)):
            b = self.batch.bvh(i)
            e = self.ctx.export(b, mesh.ntri)
            for k in ("codes", "sorted_idx", "left", "right"):
                assert np.array_equal(e[k], r["export"][k]), f"{name}: {k}"
            assert np.array_equal(_bits(e["boxes"]), _bits(r["export"]["boxes"])), f"{name}: boxes"
            info = self.ctx.bvh_info(b)
            assert info.height == r["info"].height, f"{name}: height"

    def close(self):
        self.batch.close()
        for it in self.items:
            for a in ([it["rays"]] if it["rays"] else []):
                a.free()
            for a in list((it["hits"] or {}).values()) + ([it["occ"]] if it["occ"] else []):
                a.free()
        for s in self.scenes:
            self.ctx.release(s)


@pytest.fixture(scope="module")
def setup(ctx):
    log("batch setup: %d meshes" % len(CASES))
    st = Setup(ctx)
    yield st
    st.close()


def test_batch_phase_items(setup, rq):
    ph = setup.batch.phase_items()
    ns = [m.ntri for m in setup.meshes]
    cdiv = lambda a, b: (a + b - 1) // b
    assert ph["init"] == sum(1 for n in ns if n)
    assert ph["sort0"] == sum(cdiv(n, 2048) for n in ns)
    assert ph["sort3"] == ph["sort0"]
    assert ph["karras"] == sum(cdiv(n - 1, 4096) for n in ns if n >= 2)
    # chunk-union refit in 256-leaf chunks, 4 per item; pass 2 and the
    # spanning nodes' wide records take 2048 internal nodes per item once a
    # mesh has two chunks
    assert ph["refit"] == sum(cdiv(n, 1024) for n in ns)
    assert ph["refit_top"] == sum(cdiv(n - 1, 2048) for n in ns if cdiv(n, 256) > 1)
    assert ph["widen"] == ph["refit_top"]


def test_batch_phase_times(setup, rq):
    """Device stamps of the build kernel's phases (init ... widen): one
    value per build phase, none negative, together no longer than the build
    kernel's event time."""
    ctx = setup.ctx
    ctx.profile(True)
    ctx.profile_reset()
    setup.batch.run()
    ctx.sync()
    ctx.profile(False)
    build_ms = ctx.profile_read()["batch_build"][0]
    ph = setup.batch.phase_ms()
    log(f"batch phase ms {ph}, build kernel {build_ms:.3f} ms")
    assert list(ph) == list(rq.BATCH_PHASES[:-1])
    assert all(v >= 0.0 for v in ph.values())
    assert 0.0 < sum(ph.values()) <= build_ms * 1.05 + 0.01


def test_batch_matches_per_mesh(setup):
    log("batch run (default config)")
    setup.clear()
    setup.batch.run()
    got = setup.results()
    assert setup.batch.errors() == 0
    members, grid = setup.batch.occupancy()
    log(f"batch members {members} of {grid} blocks")
    assert members == grid  # an idle device places the whole grid
    setup.check(got, "run")
    setup.check_trees()


def test_batch_against_oracle(setup, O):
    setup.clear()
    setup.batch.run()
    got = setup.results()
    for i, name in enumerate(setup.names):
        if name not in ORACLE_CASES:
            continue
        log(f"batch vs oracle {name}")
        ob = O.OracleBVH(setup.meshes[i], setup.mode)
        rays = setup.rays[i]
        if "prim" in got[i]:
            o = ob.closest(rays)
            assert np.array_equal(got[i]["prim"], o["prim"]), name
            assert np.array_equal(_bits(got[i]["t"]), _bits(o["t"])), name
        if "occ" in got[i]:
            assert np.array_equal(got[i]["occ"], ob.any(rays)["occ"]), name


def test_batch_graph_replays(setup):
    for k in range(3):
        log(f"batch graph replay {k}")
        setup.clear()
        setup.batch.run_graph()
        got = setup.results()
        assert setup.batch.errors() == 0
        setup.check(got, f"graph replay {k}")
    setup.check_trees()


@pytest.mark.parametrize("tile", [256, 1024, 8192])
def test_batch_every_config(setup, rq, tile):
    for c in rq.Batch.configs():
        log(f"batch config {c} tile={tile}")
        setup.batch.set_config(c["index"], tile, 24)
        setup.clear()
        setup.batch.run()
        got = setup.results()
        assert setup.batch.errors() == 0
        setup.check(got, f"config {c['index']} tile {tile}")
    setup.batch.set_config(0, 32768, 32)


def test_batch_every_build_kernel(setup):
    """The build kernel at each register cap (8, 7, 6, 5, 4 waves per SIMD),
    plain and replayed: the same trees and answers."""
    saved = setup.batch.build_waves()
    try:
        for waves in (8, 7, 6, 5, 4):
            log(f"batch build kernel at {waves} waves")
            assert setup.batch.build_waves(waves) == waves
            for graph in (False, True):
                setup.clear()
                setup.batch.run_graph() if graph else setup.batch.run()
                got = setup.results()
                assert setup.batch.errors() == 0
                setup.check(got, f"build waves {waves} graph {graph}")
            setup.check_trees()
    finally:
        setup.batch.build_waves(saved)


def test_batch_graph_after_queue_rebuild(setup):
    """A graph captured for one ray tile, then a plain run with another tile
    (the work queue is rebuilt), then the first configuration again: the
    replay must see a queue built for its own tile."""
    log("batch graph after work-queue rebuild")
    setup.batch.set_config(0, 2048, 16)
    setup.clear()
    setup.batch.run_graph()
    setup.check(setup.results(), "graph tile 2048")
    setup.batch.set_config(0, 32768, 16)
    setup.clear()
    setup.batch.run()
    setup.check(setup.results(), "run tile 32768")
    setup.batch.set_config(0, 2048, 16)
    setup.clear()
    setup.batch.run_graph()
    setup.check(setup.results(), "graph tile 2048 again")
    assert setup.batch.errors() == 0


def test_batch_autotune(setup):
    log("batch autotune")
    v, t, r, ms = setup.batch.autotune(1)
    assert 0 <= v < len(_rq.Batch.configs()) and t >= 256 and 1 <= r <= 64 and ms > 0
    setup.clear()
    setup.batch.run_graph()
    setup.check(setup.results(), "autotuned")
    assert setup.batch.errors() == 0


def test_batch_under_contention(setup, ctx, rq):
    """Batch runs (plain and graph) enqueued on their own stream between large
    queries on another stream, which can hold CUs when the batch starts: the
    barriers count only the blocks that started, so every run equals the
    per-mesh path with no barrier time-out, however many blocks took part."""
    mesh = rq.icosphere(6, 4, center=(0.0, 0.0, -5.0))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    n = 1 << 23
    rays = ctx.raygen_uniform(n, 11, (-1.5, -1.5, -6.5), (1.5, 1.5, -3.5))
    prim, t = rq.DeviceArray(ctx, 4 * n), rq.DeviceArray(ctx, 4 * n)
    s_q, s_b = ctx.stream_create(), ctx.stream_create()
    try:
        for k in range(4):
            log(f"batch under contention {k}")
            setup.clear()
            ctx.closest_hit_dev(b, rays, {"prim": prim, "t": t}, stream=s_q)
            if k % 2:
                setup.batch.run_graph(stream=s_b)
            else:
                setup.batch.run(stream=s_b)
            ctx.closest_hit_dev(b, rays, {"prim": prim, "t": t}, stream=s_q)
            ctx.sync()
            assert setup.batch.errors() == 0, k
            members, grid = setup.batch.occupancy()
            log(f"contended run {k}: {members} of {grid} blocks took part")
            assert 1 <= members <= grid
            setup.check(setup.results(), f"contended run {k}")
        assert ctx.device_errors() == 0
    finally:
        ctx.stream_destroy(s_q)
        ctx.stream_destroy(s_b)
        prim.free(), t.free(), rays.free()
        ctx.release(b)
        ctx.release(s)


def test_batch_ref_affine_mode(ctx):
    log("batch ref-affine mode")
    st = Setup(ctx, mode=_rq.MORTON_REF_AFFINE, cases=CASES[:6])
    try:
        st.clear()
        st.batch.run()
        st.check(st.results(), "ref-affine")
        st.check_trees()
        assert st.batch.errors() == 0
    finally:
        st.close()


def test_batch_empty(ctx, rq):
    log("batch with no meshes")
    b = rq.Batch(ctx, [])
    b.run()
    b.run_graph()
    assert b.errors() == 0
    b.close()
