"""GPU sphere primitive (the reference's scene objects): LBVH build of sphere
scenes bit-exact with the oracle, closest / any hit identical to the oracle
(ids, t, traversal counts), and the reference's own bunny scene rendered by
the GPU renderer against the colours of the reference's castRay
(tests/golden/ref_sphere_scene.npz, written from its compiled code)."""
from pathlib import Path

import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
RAYS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _cloud(rq, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-3, 3, (n, 3)).astype(np.float32) + np.float32([0, 0, -10])
    if n > 140:
        c[100:140] = c[99]  # identical spheres: ties go to the lowest index
    r = rng.uniform(0.02, 0.5, n).astype(np.float32)
    if n > 140:
        r[100:140] = r[99]
    rgb = rng.uniform(0, 1, 3 * n).astype(np.float32)
    return rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(), r, rgb)


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = np.zeros((n, 3), np.float32)
    o[: n // 5] = rng.uniform(-3, 3, (n // 5, 3)) + np.float32([0, 0, -10])  # inside the cloud
    d = rng.normal(size=(n, 3))
    d[:, 2] -= 2.5
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, np.float32),
                tmax=np.full(n, np.inf, np.float32))
    rays["tmax"][::5] = 9.5
    rays["tmin"][::11] = 8.0
    rays["tmin"][::13] = rays["tmax"][::13]
    return rays


@pytest.mark.parametrize("n", [1, 2, 3, 5000])
@pytest.mark.parametrize("mode", [0, 1])
def test_sphere_build_and_queries(ctx, O, rq, n, mode):
    log(f"spheres n={n} mode={mode}")
    sph = _cloud(rq, n, 10 + n)
    s = ctx.upload_spheres(sph)
    b = ctx.build_lbvh(s, mode)
    try:
        ob = O.OracleBVH(sph, mode)
        g, o = ctx.export(b, n), ob.arrays()
        for k in ("codes", "sorted_idx", "left", "right"):
            assert np.array_equal(g[k], o[k]), k
        assert np.array_equal(_bits(g["boxes"]), _bits(o["boxes"])), "boxes"
        rays = _rays(8000, 20 + n)
        gc, oc = ctx.closest_hit(b, rays, counted=True), ob.closest(rays)
        assert np.array_equal(gc["prim"], oc["prim"]), "closest prim"
        assert np.array_equal(_bits(gc["t"]), _bits(oc["t"])), "closest t"
        assert np.array_equal(gc["visits"], oc["visits"]) and np.array_equal(gc["tests"],
                                                                             oc["tests"])
        assert np.all(gc["u"] == 0) and np.all(gc["v"] == 0)
        g2 = ctx.closest_hit(b, rays)
        assert np.array_equal(g2["prim"], oc["prim"])
        ga, oa = ctx.any_hit(b, rays, counted=True), ob.any(rays)
        assert np.array_equal(ga["occ"], oa["occ"]) and np.array_equal(ga["visits"],
                                                                        oa["visits"])
        assert np.array_equal(ctx.any_hit(b, rays)["occ"], oa["occ"])
        # the box-filtered NONE loop of the reference (its BVH path's candidates)
        bf = O.closest_brute(sph, rays, box_filter=True)
        assert np.array_equal(gc["prim"], bf["prim"])
        if n > 100:
            assert (oc["prim"] >= 0).mean() > 0.1
    finally:
        ctx.release(b)
        ctx.release(s)


def test_reference_bunny_scene_render(ctx, O, rq):
    """The reference's scene (its bunny.obj vertices as spheres + ground) seen by
    its camera: primary rays bit-exact with the reference's render(), hit ids
    identical to its castRay, colours within 2e-5 and the image within 1/255."""
    z = np.load(GOLD / "ref_sphere_scene.npz")
    sph = rq.SphereArrays(z["cx"], z["cy"], z["cz"], z["r"], z["rgb"])
    w, h = int(z["width"]), int(z["height"])
    log(f"reference bunny scene {sph.n} spheres {w}x{h}")
    s = ctx.upload_spheres(sph)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    p = rq.render_defaults()
    p.width, p.height, p.spp = w, h, 1
    p.camera, p.shadows, p.bounce = rq.CAMERA_REF_JITTER, 0, 0
    r = rq.Renderer(ctx, p)
    try:
        r.frame(b)
        ctx.sync()
        assert ctx.device_errors() == 0
        prim = r.rays(0).download()
        for k in ("dx", "dy", "dz"):
            assert np.array_equal(_bits(prim[k]), _bits(z[k])), k
        g = r.samples()
        assert np.array_equal(g["hit0"], z["hit"]), "hit ids vs the reference's castRay"
        assert np.all(g["shadow0"] == 0)
        assert np.max(np.abs(g["rgb"] - z["colors"])) < 2e-5
        _, ppm = r.image()
        ref_ppm = O.resolve_ppm(z["colors"], w * h, 1)
        diff = np.abs(ppm.astype(np.int16) - ref_ppm.astype(np.int16))
        assert diff.max() <= 1, f"image differs by {diff.max()}/255"
        cnt = r.counts()
        assert cnt.primary == w * h and cnt.shadow == 0 and cnt.bounce == 0
    finally:
        r.close()
        ctx.release(b)
        ctx.release(s)


def test_batch_refuses_sphere_scenes(ctx, rq):
    sph = _cloud(rq, 10, 1)
    s = ctx.upload_spheres(sph)
    try:
        with pytest.raises(rq.RQError):
            rq.Batch(ctx, [dict(scene=s, rays=None, hits=None, occ=None)])
    finally:
        ctx.release(s)


@pytest.mark.parametrize("variant", ["a", "b"])
def test_reference_scene_materials_render(ctx, O, rq, variant):
    """castRay's material switch on the reference's bunny scene (fixture from its
    compiled castRay with the materials set): refraction rays traced in the
    second batch, hit ids identical to the oracle's at both depths, colours
    within 2e-5 of the reference's (powf of the specular term is not correctly
    rounded on either side), bit-exact elsewhere up to the few box-filtered
    silhouette samples."""
    z = np.load(GOLD / "ref_sphere_scene.npz")
    sph = rq.SphereArrays(z["cx"], z["cy"], z["cz"], z["r"], z["rgb"])
    mat = z[f"mat_{variant}"]
    w, h = int(z["width"]), int(z["height"])
    log(f"reference bunny scene materials {variant}")
    s = ctx.upload_spheres(sph)
    ctx.set_materials(s, mat)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    p = rq.render_defaults()
    p.width, p.height, p.spp = w, h, 1
    p.camera, p.shadows, p.bounce, p.materials = rq.CAMERA_REF_JITTER, 0, 0, 1
    r = rq.Renderer(ctx, p)
    try:
        r.frame(b)
        ctx.sync()
        assert ctx.device_errors() == 0
        g = r.samples()
        rays = O.camera_rays_ref(w, h, 1, O.mt19937_jitter(2 * w * h))
        o = O.render_samples(O.OracleBVH(sph, 1), sph, rays,
                             O.shade_params(shadows=0, prim_rgb=sph.rgb, prim_mat=mat))
        assert np.array_equal(g["hit0"], o["hit0"])
        assert np.array_equal(g["hit1"], o["hit1"])
        assert (g["hit1"] >= 0).sum() > 100
        col, ref = g["rgb"].reshape(-1, 3), z[f"colors_mat_{variant}"].reshape(-1, 3)
        same = np.all(_bits(o["rgb"].reshape(-1, 3)) == _bits(ref), axis=1)
        assert (~same).sum() <= 16
        assert np.max(np.abs(col[same] - ref[same])) < 2e-5
        cnt = r.counts()
        assert cnt.primary == w * h and cnt.shadow == 0
        assert cnt.bounce == int((mat[o["hit0"][o["hit0"] >= 0]] == 1).sum())
    finally:
        r.close()
        ctx.release(b)
        ctx.release(s)


def test_reference_scene_kd_occlusion_render(ctx, rq):
    """castRay's KDTREE case: the GPU any hit of the primary rays against the
    reference's own kd-tree (kdtreeIntersect, fixture): identical per ray."""
    z = np.load(GOLD / "ref_sphere_scene.npz")
    sph = rq.SphereArrays(z["cx"], z["cy"], z["cz"], z["r"], z["rgb"])
    w, h = int(z["width"]), int(z["height"])
    s = ctx.upload_spheres(sph)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    p = rq.render_defaults()
    p.width, p.height, p.spp = w, h, 1
    p.camera, p.shadows, p.mode = rq.CAMERA_REF_JITTER, 0, rq.RENDER_OCCLUSION
    r = rq.Renderer(ctx, p)
    try:
        r.frame(b)
        ctx.sync()
        assert ctx.device_errors() == 0
        g = r.samples()
        assert np.array_equal(g["shadow0"], z["kd_any"])
        assert np.all(g["hit0"] == -1)
        col = g["rgb"].reshape(-1, 3)
        assert np.all(col[z["kd_any"] == 1] == 0)
        assert np.all(col[z["kd_any"] == 0] == np.float32([0.6, 0.8, 1.0]))
        cnt = r.counts()
        assert cnt.primary == w * h and cnt.bounce == 0 and cnt.shadow == 0
    finally:
        r.close()
        ctx.release(b)
        ctx.release(s)


def test_render_params_refused(ctx, rq):
    p = rq.render_defaults()
    p.materials, p.bounce = 1, 1
    with pytest.raises(rq.RQError):
        rq.Renderer(ctx, p)
    p = rq.render_defaults()
    p.mode = 7
    with pytest.raises(rq.RQError):
        rq.Renderer(ctx, p)
