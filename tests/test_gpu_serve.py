"""Serving loop: results of every request are bit-identical when it runs in a
mix (continuous batching with other requests, chunked over several trace
launches, new meshes built meanwhile) and when it runs alone, and equal to
the plain path (rq_build_lbvh + rq_closest_hit / rq_any_hit) on the same mesh
and rays; admission respects the arena; oversize requests are refused."""
import numpy as np
import pytest

import rq as _rq
from conftest import log
from helpers import random_rays, random_soup

pytestmark = pytest.mark.gpu

SMALL = dict(mesh_slots=8, slot_tris=1 << 16, arena_bytes=256 << 20, max_inflight=16,
             step_rays=1 << 16, chunk_rays=1 << 14, ray_tile=1024, hbm_budget=8 << 30)


def _meshes():
    return [
        _rq.icosphere(4, 11, center=(0, 0, -5)),
        random_soup(20000, 21, spread=1.5, size=0.02),
        _rq.icosphere(5, 12, center=(0.2, 0.1, -5)),
        random_soup(3000, 22, spread=1.5, size=0.05),
    ]


def _new_meshes():
    return [_rq.icosphere(3, 13, center=(0, 0, -5)), random_soup(7000, 23, spread=1.5, size=0.03)]


def _requests(slots, new):
    cam = dict(eye=(0.0, 0.0, 0.0), look_at=(0.0, 0.0, -5.0), up=(0.0, 1.0, 0.0), fov_deg=35.0,
               width=200)
    box = ((-2.0, -2.0, -7.0), (2.0, 2.0, -3.0))
    Q, A, B = _rq.Q_CLOSEST, _rq.Q_ANY, _rq.Q_CLOSEST | _rq.Q_ANY
    return [
        dict(mesh=slots[0], nrays=40000, query=B, pinhole=cam, seed=1),
        dict(mesh=slots[1], nrays=5000, query=Q, rays=random_rays(5000, 31)),
        dict(mesh=slots[2], nrays=70001, query=B, box=box, seed=2),
        dict(mesh=slots[3], nrays=1, query=Q, box=box, seed=3),
        dict(new_mesh=new[0], nrays=30000, query=B, pinhole=cam, seed=4),
        dict(mesh=slots[0], nrays=100, query=A, rays=random_rays(100, 32)),
        dict(mesh=slots[2], nrays=16384, query=Q, pinhole=dict(cam, width=128), seed=5),
        dict(mesh=slots[1], nrays=50000, query=A, box=box, seed=6),
        dict(new_mesh=new[1], nrays=12345, query=B, box=box, seed=7),
        dict(mesh=slots[3], nrays=20000, query=B, rays=random_rays(20000, 33)),
        dict(mesh=slots[2], nrays=999, query=B, pinhole=cam, seed=8),
        dict(mesh=slots[1], nrays=65536, query=Q, pinhole=cam, seed=9),
    ]


def _submit(srv, spec, pinned):
    n = spec["nrays"]
    q = spec["query"]
    if pinned:
        outs = dict(prim=_rq.HostArray(n, np.int32), t=_rq.HostArray(n, np.float32),
                    occ=_rq.HostArray(n, np.uint8))
        arrs = {k: v.array for k, v in outs.items()}
    else:
        outs = None
        arrs = dict(prim=np.empty(n, np.int32), t=np.empty(n, np.float32),
                    occ=np.empty(n, np.uint8))
    for a in arrs.values():
        a.fill(0xA5 if a.dtype != np.uint8 else 0x7F)
    kw = {k: spec[k] for k in ("mesh", "new_mesh", "rays", "pinhole", "box", "seed") if k in spec}
    tk = srv.submit(n, q, prim=arrs["prim"] if q & _rq.Q_CLOSEST else None,
                    t=arrs["t"] if q & _rq.Q_CLOSEST else None,
                    occ=arrs["occ"] if q & _rq.Q_ANY else None, **kw)
    return tk, arrs, outs


def _result(spec, arrs):
    out = {}
    if spec["query"] & _rq.Q_CLOSEST:
        out["prim"] = arrs["prim"].copy()
        out["t"] = arrs["t"].copy().view(np.uint32)
    if spec["query"] & _rq.Q_ANY:
        out["occ"] = arrs["occ"].copy()
    return out


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{what}: {k}"


@pytest.fixture(scope="module")
def served():
    """The mix: every request submitted at once, then each alone."""
    log("serve: create server")
    srv = _rq.Server(0, **SMALL)
    meshes = _meshes()
    slots = [srv.load_mesh(m) for m in meshes]
    new = _new_meshes()
    specs = _requests(slots, new)
    log(f"serve: mix of {len(specs)} requests")
    t0 = _rq.server_clock()
    sub = [_submit(srv, s, pinned=(i % 2 == 0)) for i, s in enumerate(specs)]
    status = [srv.wait(tk) for tk, _, _ in sub]
    mix = [_result(s, a) for s, (_, a, _) in zip(specs, sub)]
    stats_mix = srv.stats()
    alone, alone_status = [], []
    for i, s in enumerate(specs):
        log(f"serve: alone {i}")
        tk, arrs, outs = _submit(srv, s, pinned=(i % 2 == 1))
        alone_status.append(srv.wait(tk))
        alone.append(_result(s, arrs))
    yield dict(srv=srv, meshes=meshes, new=new, slots=slots, specs=specs, status=status, mix=mix,
               alone=alone, alone_status=alone_status, stats=stats_mix, t0=t0)
    srv.close()


def test_serve_all_requests_complete(served):
    for st in served["status"] + served["alone_status"]:
        assert st.state == _rq.REQ_DONE, st.state
        assert served["t0"] <= st.t_submit <= st.t_admit <= st.t_done
    s = served["stats"]
    assert s["failed"] == 0 and s["completed"] >= len(served["specs"])
    assert s["builds"] >= len(served["meshes"]) + len(served["new"])


def test_serve_continuous_batching_happened(served):
    s = served["stats"]
    assert s["max_step_requests"] >= 2, "no trace launch held more than one request"
    big = [st.steps for sp, st in zip(served["specs"], served["status"]) if sp["nrays"] > 16384]
    assert big and min(big) >= 2, "a request larger than its chunk must span several launches"


def test_serve_mix_equals_alone(served):
    for i, (a, b) in enumerate(zip(served["mix"], served["alone"])):
        _same(a, b, f"request {i} mix vs alone")


def test_serve_equals_plain_path(served, ctx):
    for i, spec in enumerate(served["specs"]):
        mesh = spec["new_mesh"] if "new_mesh" in spec else served["meshes"][
            served["slots"].index(spec["mesh"])]
        n = spec["nrays"]
        log(f"serve: plain path {i}")
        if "rays" in spec:
            rays = _rq.DeviceRays.from_host(ctx, spec["rays"])
        elif "pinhole" in spec:
            p = spec["pinhole"]
            rays = ctx.raygen_pinhole(n, spec["seed"], p["eye"], p["look_at"], p["up"],
                                      p["fov_deg"], p["width"])
        else:
            rays = ctx.raygen_uniform(n, spec["seed"], *spec["box"])
        s = ctx.upload_scene(mesh)
        b = ctx.build_lbvh(s, _rq.MORTON_SCENE_AABB)
        host = rays.download()
        ref = {}
        if spec["query"] & _rq.Q_CLOSEST:
            h = ctx.closest_hit(b, host)
            ref["prim"] = h["prim"]
            ref["t"] = h["t"].view(np.uint32)
        if spec["query"] & _rq.Q_ANY:
            ref["occ"] = ctx.any_hit(b, host)["occ"]
        ctx.release(b)
        ctx.release(s)
        rays.free()
        _same(served["mix"][i], ref, f"request {i} vs plain path")
        if "prim" in ref:
            assert (ref["prim"] >= 0).any() or n < 10


def test_serve_admission_and_refusals(rq):
    log("serve: small arena")
    cfg = dict(SMALL, arena_bytes=2 << 20, max_inflight=4, mesh_slots=2)
    with rq.Server(0, **cfg) as srv:
        slot = srv.load_mesh(rq.icosphere(3, 5, center=(0, 0, -5)))
        box = ((-2.0, -2.0, -7.0), (2.0, 2.0, -3.0))
        tks, outs = [], []
        for k in range(20):
            prim, occ = np.empty(10000, np.int32), np.empty(10000, np.uint8)
            tks.append(srv.submit(10000, rq.Q_CLOSEST | rq.Q_ANY, prim=prim, occ=occ, mesh=slot,
                                  box=box, seed=100 + k))
            outs.append((prim, occ))
        sts = [srv.wait(t) for t in tks]
        assert all(s.state == rq.REQ_DONE for s in sts)
        st = srv.stats()
        assert 0 < st["arena_peak"] <= 2 << 20
        # every result equals the same request alone
        for k in (0, 7, 19):
            prim, occ = np.empty(10000, np.int32), np.empty(10000, np.uint8)
            t = srv.submit(10000, rq.Q_CLOSEST | rq.Q_ANY, prim=prim, occ=occ, mesh=slot, box=box,
                           seed=100 + k)
            srv.wait(t)
            assert np.array_equal(prim, outs[k][0]) and np.array_equal(occ, outs[k][1])
        with pytest.raises(rq.NoMemory):
            big = np.empty(100000, np.int32)
            srv.submit(100000, rq.Q_CLOSEST, prim=big, mesh=slot, box=box)
        with pytest.raises(rq.NoMemory):
            srv.load_mesh(random_soup(70000, 3))
        with pytest.raises(rq.RQError):
            srv.submit(10, rq.Q_CLOSEST, prim=np.empty(10, np.int32), mesh=1, box=box)


def test_serve_resident_slots_are_never_recycled(rq):
    """Meshes of rq_server_load_mesh keep their slots: uploads recycle only
    the slots beyond them, a request may not name an upload's slot, and a
    server whose slots all hold resident meshes refuses uploads."""
    log("serve: resident slots")
    cfg = dict(SMALL, mesh_slots=3)
    box = ((-2.0, -2.0, -7.0), (2.0, 2.0, -3.0))
    with rq.Server(0, **cfg) as srv:
        meshes = [rq.icosphere(3, 40 + k, center=(0, 0, -5)) for k in range(2)]
        slots = [srv.load_mesh(m) for m in meshes]
        ref = []
        for sl in slots:
            prim = np.empty(5000, np.int32)
            srv.wait(srv.submit(5000, rq.Q_CLOSEST, prim=prim, mesh=sl, box=box, seed=7))
            ref.append(prim)
        up_slots = set()
        for k in range(4):  # uploads through the one free slot, none evicts a resident mesh
            prim = np.empty(3000, np.int32)
            st = srv.wait(srv.submit(3000, rq.Q_CLOSEST, prim=prim,
                                     new_mesh=random_soup(2000 + k, 60 + k, spread=1.5),
                                     box=box, seed=k))
            assert st.state == rq.REQ_DONE
            up_slots.add(st.mesh)
        assert len(up_slots) == 1 and not up_slots & set(slots)
        for sl, r in zip(slots, ref):
            prim = np.empty(5000, np.int32)
            srv.wait(srv.submit(5000, rq.Q_CLOSEST, prim=prim, mesh=sl, box=box, seed=7))
            assert np.array_equal(prim, r)
        with pytest.raises(rq.RQError) as e:  # an upload's slot is not a resident mesh
            srv.submit(10, rq.Q_CLOSEST, prim=np.empty(10, np.int32), mesh=up_slots.pop(),
                       box=box)
        assert e.value.code == rq.RQ_E_INVALID
        srv.load_mesh(rq.icosphere(2, 3, center=(0, 0, -5)))  # the last slot becomes resident
        with pytest.raises(rq.NoMemory):
            srv.submit(10, rq.Q_CLOSEST, prim=np.empty(10, np.int32),
                       new_mesh=random_soup(100, 1), box=box)
        with pytest.raises(rq.NoMemory):
            srv.load_mesh(rq.icosphere(2, 4, center=(0, 0, -5)))


@pytest.mark.parametrize("pinned", [False, True])
def test_serve_cancel(rq, pinned):
    """Cancelled requests finish as RQ_E_CANCELED, queued ones at once and a
    running one once its launches retire; the others complete with their
    alone results, and the server keeps serving. Page-locked outputs take each
    chunk's result copy at its launch, pageable ones one copy at the end."""
    log(f"serve: cancel, page-locked outputs {pinned}")
    cfg = dict(SMALL, max_inflight=2, chunk_rays=1 << 12, step_rays=1 << 13)
    box = ((-2.0, -2.0, -7.0), (2.0, 2.0, -3.0))
    with rq.Server(0, **cfg) as srv:
        slot = srv.load_mesh(rq.icosphere(5, 9, center=(0, 0, -5)))
        n = 60000
        keep = [rq.HostArray(n, np.int32) for _ in range(8)] if pinned else []
        outs = [h.array for h in keep] if pinned else [np.empty(n, np.int32) for _ in range(8)]
        for o in outs:
            o.fill(-7)
        tks = [srv.submit(n, rq.Q_CLOSEST, prim=outs[k], mesh=slot, box=box, seed=200 + k)
               for k in range(8)]
        cancelled = {1, 5, 6}
        for k in sorted(cancelled):
            srv.cancel(tks[k])
        st = [srv.wait(t) for t in tks]
        for k in range(8):
            if k in cancelled:
                assert st[k].state in (rq.RQ_E_CANCELED, rq.REQ_DONE), (k, st[k].state)
            else:
                assert st[k].state == rq.REQ_DONE, (k, st[k].state)
        assert st[5].state == rq.RQ_E_CANCELED and st[6].state == rq.RQ_E_CANCELED  # queued
        with pytest.raises(rq.RQError):
            srv.cancel(tks[0])  # finished
        for k in (0, 2, 7):
            prim = np.empty(n, np.int32)
            srv.wait(srv.submit(n, rq.Q_CLOSEST, prim=prim, mesh=slot, box=box, seed=200 + k))
            assert np.array_equal(prim, outs[k])
        s = srv.stats()
        assert s["failed"] >= 2
