"""CPU: the sphere primitive, i.e. the reference's own scene objects
(accelerators.h:54-92, createScene_new main.cpp:599-721, castRay's NONE loop
main.cpp:376-386 and its DIFFUSE_AND_GLOSSY shading main.cpp:447-492).

The oracle's sphere path is pinned by the reference's compiled code: its NONE
loop hits (ref_vectors.npz, live oracle/_ref when present), its castRay
colours on its own bunny scene (ref_sphere_scene.npz), and brute force for
the oracle's LBVH traversal."""
from pathlib import Path

import numpy as np
import pytest

from conftest import REFERENCE

GOLD = Path(__file__).resolve().parent / "golden"


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _spheres(rq, c, r, rgb=None):
    c = np.asarray(c, np.float32)
    n = len(c)
    rgb = np.tile(np.float32([0.8, 0.7, 0.0]), n) if rgb is None else np.asarray(rgb, np.float32)
    return rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(),
                           np.asarray(r, np.float32).copy(), rgb)


def _fixture_spheres(rq, z):
    return rq.SphereArrays(z["cx"], z["cy"], z["cz"], z["r"], z["rgb"])


def test_reference_scene_loader_semantics(rq, tmp_path):
    p = tmp_path / "a.obj"
    # reading stops at the first token that is not `v` (main.cpp:694-697)
    p.write_text("v 0.1 0.2 0.3\nv -1 2.5 0\nf 1 2 3\nv 9 9 9\n")
    s = rq.reference_scene(p, clones=2)
    assert s.n == 2 * 2 + 1
    v = np.float32([[0.1, 0.2, 0.3], [-1, 2.5, 0]])
    for clone in range(2):
        shift = np.float32(clone * 20)
        c = v * np.float32(100) + shift
        c[:, 1] = c[:, 1] + np.float32(-10)
        c[:, 2] = c[:, 2] + np.float32(-60)
        k = slice(2 * clone, 2 * clone + 2)
        assert np.array_equal(_bits(s.cx[k]), _bits(c[:, 0]))
        assert np.array_equal(_bits(s.cy[k]), _bits(c[:, 1]))
        assert np.array_equal(_bits(s.cz[k]), _bits(c[:, 2]))
    assert np.all(s.r[:-1] == np.float32(0.05)) and s.r[-1] == 100
    assert np.array_equal(s.rgb[:-3], np.tile(np.float32([0.8, 0.7, 0.0]), 4))
    assert (s.cx[-1], s.cy[-1], s.cz[-1]) == (np.float32(0.93591022), np.float32(-105.47120094),
                                              np.float32(-43.2363205))
    assert rq.reference_scene(p, 1, 1).n == 2
    q = tmp_path / "b.obj"
    q.write_text("# comment first\nv 1 2 3\n")
    assert rq.reference_scene(q).n == 1  # only the ground sphere
    with pytest.raises(Exception):
        rq.reference_scene(tmp_path / "missing.obj")


@pytest.mark.skipif(not (REFERENCE / "models" / "bunny.obj").exists(),
                    reason="reference models not present")
def test_reference_scene_loader_on_bunny(rq):
    s = rq.reference_scene(REFERENCE / "models" / "bunny.obj")
    rows = [l.split()[1:4] for l in (REFERENCE / "models" / "bunny.obj").read_text().splitlines()
            if l.startswith("v ")]
    v = np.array(rows, np.float64).astype(np.float32)
    assert s.n == len(v) + 1 == 35948
    assert np.array_equal(_bits(s.cx[:-1]), _bits(v[:, 0] * np.float32(100) + np.float32(0)))
    assert np.array_equal(_bits(s.cz[:-1]),
                          _bits((v[:, 2] * np.float32(100) + np.float32(0)) + np.float32(-60)))


def test_oracle_sphere_none_loop_equals_reference(O, rq):
    """The oracle's brute force is the reference's NONE loop (fixture from its
    compiled code: Sphere::raySphereIntersect + castRay's t0 / t1 rule)."""
    z = np.load(GOLD / "ref_vectors.npz")
    sph = _spheres(rq, z["sphere_centres"], z["sphere_radii"])
    o, d = z["sphere_ray_o"], z["sphere_ray_d"]
    n = len(o)
    rays = dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, np.float32),
                tmax=np.full(n, np.inf, np.float32))
    b = O.closest_brute(sph, rays)
    assert np.array_equal(b["prim"], z["sphere_hit_none"])
    hit = b["prim"] >= 0
    assert np.array_equal(_bits(b["t"][hit]), _bits(z["sphere_t_none"][hit]))
    assert hit.mean() > 0.2


def _sphere_rays(n, seed, inside_frac=0.2):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-3, 3, (n, 3)).astype(np.float32) + np.float32([0, 0, -10])
    o[: int(n * (1 - inside_frac))] = 0.0
    d = rng.normal(size=(n, 3))
    d[:, 2] -= 2.5
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, np.float32),
                tmax=np.full(n, np.inf, np.float32))
    rays["tmax"][::5] = 9.5       # clipped
    rays["tmin"][::11] = 8.0      # starts late
    rays["tmin"][::13] = rays["tmax"][::13]  # inactive
    return rays


@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_sphere_bvh_equals_brute(O, rq, mode):
    rng = np.random.default_rng(5 + mode)
    n = 3000
    c = rng.uniform(-3, 3, (n, 3)).astype(np.float32) + np.float32([0, 0, -10])
    c[100:140] = c[99]  # identical spheres: ties go to the lowest index
    r = rng.uniform(0.02, 0.5, n).astype(np.float32)
    r[100:140] = r[99]
    sph = _spheres(rq, c, r)
    rays = _sphere_rays(6000, 9 + mode)
    bvh = O.OracleBVH(sph, mode)
    got, ref = bvh.closest(rays), O.closest_brute(sph, rays, box_filter=True)
    assert np.array_equal(got["prim"], ref["prim"])
    assert np.array_equal(_bits(got["t"]), _bits(ref["t"]))
    assert (ref["prim"] >= 0).mean() > 0.1
    none = O.closest_brute(sph, rays)  # unfiltered NONE loop: differs at most on a few rays
    assert (none["prim"] != ref["prim"]).mean() < 2e-3
    occ = bvh.any(rays)["occ"]
    assert np.array_equal(occ[ref["prim"] >= 0], np.ones((ref["prim"] >= 0).sum(), np.uint8))
    assert np.all(occ <= O.any_brute(sph, rays))
    # boxes: leaves are centre -/+ r, the root holds everything
    a = bvh.arrays()
    leaf = a["boxes"][n - 1:]
    p = a["sorted_idx"]
    assert np.array_equal(_bits(leaf[:, 0]), _bits((c[p, 0] - r[p]) + np.float32(0)))
    assert np.array_equal(_bits(leaf[:, 5]), _bits((c[p, 2] + r[p]) + np.float32(0)))
    assert np.all(a["boxes"][0, :3] <= leaf[:, :3].min(axis=0))


def test_oracle_sphere_small_counts(O, rq):
    for n in (1, 2, 3):
        c = np.float32([[0, 0, -5], [0.3, 0, -6], [0, 0.2, -7]])[:n]
        sph = _spheres(rq, c, np.float32([0.5, 0.5, 0.5])[:n])
        rays = _sphere_rays(500, 3)
        got, ref = O.OracleBVH(sph, 1).closest(rays), O.closest_brute(sph, rays, box_filter=True)
        assert np.array_equal(got["prim"], ref["prim"])
        assert np.array_equal(_bits(got["t"]), _bits(ref["t"]))


def test_oracle_sphere_scene_render_equals_reference_castray(O, rq):
    """castRay's colours on the reference's bunny scene (fixture written from
    the reference's compiled code): the oracle's sphere render, shadows off as
    the reference's trace_more stub, matches colour for colour."""
    z = np.load(GOLD / "ref_sphere_scene.npz")
    sph = _fixture_spheres(rq, z)
    w, h = int(z["width"]), int(z["height"])
    rays = O.camera_rays_ref(w, h, 1, O.mt19937_jitter(2 * w * h))
    # the oracle camera is the reference's render() camera, bit for bit
    for k in ("dx", "dy", "dz"):
        assert np.array_equal(_bits(rays[k]), _bits(z[k])), k
    bvh = O.OracleBVH(sph, 1)
    sp = O.shade_params(shadows=0, prim_rgb=sph.rgb)
    out = O.render_samples(bvh, sph, rays, sp)
    # the oracle's NONE loop is the reference's, ray for ray
    none = O.closest_brute(sph, rays)
    assert np.array_equal(none["prim"], z["hit"])
    hit = z["hit"] >= 0
    assert np.array_equal(_bits(none["t"][hit]), _bits(z["t"][hit]))
    # a box hierarchy sees the box-filtered NONE loop (the reference BVH path's
    # candidates): it differs on the few rays whose float sphere test accepts a
    # sphere whose box the ray misses (d2 cancellation at distance ~57)
    filt = O.closest_brute(sph, rays, box_filter=True)
    assert np.array_equal(out["hit0"], filt["prim"])
    same = filt["prim"] == z["hit"]
    assert (~same).sum() <= 16, (~same).sum()
    col = out["rgb"].reshape(-1, 3)
    ref = z["colors"].reshape(-1, 3)
    assert np.array_equal(_bits(col[same]), _bits(ref[same])), np.abs(col - ref)[same].max()
    bunny = (z["hit"] >= 0) & (z["hit"] < sph.n - 1)
    assert bunny.sum() > 100 and (z["hit"] == sph.n - 1).sum() > 1000


def test_oracle_sphere_render_live_reference(O, rq):
    """Same check against the live reference build on a random sphere cloud."""
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(77)
    n = 700
    c = rng.uniform(-6, 6, (n, 3)).astype(np.float32) + np.float32([0, 0, -30])
    r = rng.uniform(0.1, 0.9, n).astype(np.float32)
    rgb = rng.uniform(0, 1, 3 * n).astype(np.float32)
    sph = _spheres(rq, c, r, rgb)
    rays = O.camera_rays_ref(64, 48, 1, O.mt19937_jitter(2 * 64 * 48))
    ref = O.ref_sphere_scene_colors(sph, rays)
    out = O.render_samples(O.OracleBVH(sph, 1), sph, rays,
                           O.shade_params(shadows=0, prim_rgb=rgb))
    same = O.closest_brute(sph, rays, box_filter=True)["prim"] == O.closest_brute(sph, rays)["prim"]
    assert same.mean() > 0.998
    col, refc = out["rgb"].reshape(-1, 3), ref.reshape(-1, 3)
    assert np.array_equal(_bits(col[same]), _bits(refc[same]))
    assert (out["hit0"] >= 0).mean() > 0.2


@pytest.mark.parametrize("variant", ["a", "b"])
def test_oracle_sphere_materials_equal_reference_castray(O, rq, variant):
    """castRay's material switch (main.cpp:417-446) with the reference's own
    compiled code on its bunny scene: REFLECTION_AND_REFRACTION traces the
    refraction ray (ior 3) weighted by 1 - fresnel (ior 2), REFLECTION gives the
    grey 1 - kr, depth 3 is the sky. Variant b refracts into the ground sphere,
    so its refraction rays start inside it. Colour for colour."""
    z = np.load(GOLD / "ref_sphere_scene.npz")
    sph = _fixture_spheres(rq, z)
    mat = z[f"mat_{variant}"]
    w, h = int(z["width"]), int(z["height"])
    rays = O.camera_rays_ref(w, h, 1, O.mt19937_jitter(2 * w * h))
    out = O.render_samples(O.OracleBVH(sph, 1), sph, rays,
                           O.shade_params(shadows=0, prim_rgb=sph.rgb, prim_mat=mat))
    col, ref = out["rgb"].reshape(-1, 3), z[f"colors_mat_{variant}"].reshape(-1, 3)
    same = np.all(_bits(col) == _bits(ref), axis=1)
    # box-filtered silhouette hits (see the plain scene test) may differ on a handful
    assert (~same).sum() <= 16, (~same).sum()
    h0 = out["hit0"]
    assert (mat[h0[h0 >= 0]] == 1).sum() > 100 and (mat[h0[h0 >= 0]] == 2).sum() > 100
    assert (out["hit1"] >= 0).sum() > 100  # refraction rays that hit something
    if variant == "b":
        assert (out["hit1"] == sph.n - 1).sum() > 1000  # inside the ground sphere


def test_oracle_kd_mode_equals_reference_kdtree(O, rq):
    """castRay's KDTREE case (main.cpp:362-370): the reference's own SAH kd-tree
    and kdtreeIntersect (fixture from its compiled code) against the oracle's
    any hit of the primary rays: black on a hit, sky otherwise."""
    z = np.load(GOLD / "ref_sphere_scene.npz")
    sph = _fixture_spheres(rq, z)
    w, h = int(z["width"]), int(z["height"])
    rays = O.camera_rays_ref(w, h, 1, O.mt19937_jitter(2 * w * h))
    out = O.render_samples(O.OracleBVH(sph, 1), sph, rays, O.shade_params(shadows=0, mode=1))
    assert np.array_equal(out["shadow0"], z["kd_any"])
    assert 0.1 < z["kd_any"].mean() < 0.9
    col = out["rgb"].reshape(-1, 3)
    assert np.all(col[z["kd_any"] == 1] == 0)
    assert np.all(col[z["kd_any"] == 0] == np.float32([0.6, 0.8, 1.0]))
    assert np.all(out["hit0"] == -1)
    # the KD path's leaf test has no t range: any sphere with its centre ahead
    # and the ray within its radius, i.e. the brute-force any hit from t = 0
    assert np.array_equal(O.any_brute(sph, rays), z["kd_any"])


def test_oracle_materials_and_kd_live_reference(O, rq):
    """The same two checks against the live reference build on a random cloud."""
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(78)
    n = 600
    c = rng.uniform(-6, 6, (n, 3)).astype(np.float32) + np.float32([0, 0, -30])
    r = rng.uniform(0.2, 1.5, n).astype(np.float32)
    rgb = rng.uniform(0, 1, 3 * n).astype(np.float32)
    sph = _spheres(rq, c, r, rgb)
    mat = rng.integers(0, 3, n).astype(np.uint8)
    rays = O.camera_rays_ref(64, 48, 1, O.mt19937_jitter(2 * 64 * 48))
    ref = O.ref_sphere_scene_colors(sph, rays, mat).reshape(-1, 3)
    out = O.render_samples(O.OracleBVH(sph, 1), sph, rays,
                           O.shade_params(shadows=0, prim_rgb=rgb, prim_mat=mat))
    same = np.all(_bits(out["rgb"].reshape(-1, 3)) == _bits(ref), axis=1)
    assert same.mean() > 0.995, same.mean()
    assert (out["hit1"] >= 0).sum() > 50
    kd = O.ref_kdtree_any(sph, rays)
    occ = O.render_samples(O.OracleBVH(sph, 1), sph, rays, O.shade_params(shadows=0, mode=1))
    assert (occ["shadow0"] == kd).mean() > 0.998
