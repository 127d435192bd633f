"""CPU: the oracle's shading (castRay's DIFFUSE_AND_GLOSSY branch,
main.cpp:447-492, with a real shadow ray and the fresnel-weighted mirror
bounce) on closed-form single-light cases."""
import numpy as np

from helpers import RAY_FIELDS, mesh_from

SKY = np.array([0.6, 0.8, 1.0])
KD = np.array([0.815, 0.235, 0.031]) * 0.8  # getDiffuseColor(0.2) * 0.8


def _one_ray(o, d):
    a = np.array([[*o, *d, 0.0, np.inf]], np.float32)
    return {k: a[:, i].copy() for i, k in enumerate(RAY_FIELDS)}


def _floor(extra=()):
    v = [(-10, -10, -5), (10, -10, -5), (0, 10, -5), *extra]
    return mesh_from(v, np.arange(len(v)))


def test_head_on_lambert_and_specular(O):
    """Triangle facing +z, light straight above the hit point: LdotN = 1 and
    the mirror of the light direction points at the eye, so the colour is
    (E*1*kd)/2 + 0.5*E*1^25."""
    mesh = _floor()
    b = O.OracleBVH(mesh, 1)
    out = O.render_samples(b, mesh, _one_ray((0, 0, 0), (0, 0, -1)), O.shade_params(light=(0, 0, 30)))
    assert out["hit0"][0] == 0 and out["shadow0"][0] == 0
    assert np.allclose(out["rgb"], KD / 2 + 0.5, atol=1e-6)
    assert out["counts"][0] == 2  # primary + shadow


def test_shadowed_keeps_specular(O):
    """An occluder between hit and light zeroes the Lambert term only; the
    reference adds the specular term unshadowed (main.cpp:473-475)."""
    mesh = _floor(extra=[(-1, -1, 10), (1, -1, 10), (0, 1, 10)])
    b = O.OracleBVH(mesh, 1)
    out = O.render_samples(b, mesh, _one_ray((0, 0, 0), (0, 0, -1)), O.shade_params(light=(0, 0, 30)))
    assert out["hit0"][0] == 0 and out["shadow0"][0] == 1
    assert np.allclose(out["rgb"], [0.5, 0.5, 0.5], atol=1e-6)


def test_surface_colour_added_per_light(O):
    mesh = _floor()
    b = O.OracleBVH(mesh, 1)
    out = O.render_samples(b, mesh, _one_ray((0, 0, 0), (0, 0, -1)),
                           O.shade_params(light=(0, 0, 30), surface=(0.1, 0.2, 0.3)))
    assert np.allclose(out["rgb"], KD / 2 + 0.5 + np.array([0.1, 0.2, 0.3]), atol=1e-6)


def test_miss_is_sky(O):
    mesh = _floor()
    b = O.OracleBVH(mesh, 1)
    out = O.render_samples(b, mesh, _one_ray((0, 0, 0), (0, 0, 1)), O.shade_params(light=(0, 0, 30)))
    assert out["hit0"][0] == -1 and np.allclose(out["rgb"], SKY)


def test_bounce_fresnel_normal_incidence(O):
    """Normal incidence, ior 2: kr = ((1/3)^2 + (1/3)^2) / 2 = 1/9; the mirror
    ray leaves the scene and sees the sky."""
    mesh = _floor()
    b = O.OracleBVH(mesh, 1)
    out = O.render_samples(b, mesh, _one_ray((0, 0, 0), (0, 0, -1)), O.shade_params(light=(0, 0, 30), bounce=1))
    assert out["hit1"][0] == -1
    assert np.allclose(out["rgb"], KD / 2 + 0.5 + SKY / 9, atol=1e-6)
    assert out["counts"][0] == 3  # primary + shadow + bounce


def test_resolve_ppm_truncates(O):
    rgb = np.array([0.5, 1.2, 0.0, 0.5, 0.2, 0.004], np.float32)  # one pixel, 2 samples
    assert O.resolve_ppm(rgb, 1, 2).tolist() == [127, 178, 0]
