"""CPU: the oracle's Morton codes against the reference's own expandBits /
morton3D (compiled from its sources into oracle/_ref) and closed forms."""
import numpy as np
import pytest

from helpers import mesh_from


def _spread_bits_independent(v: int) -> int:
    # bit k of v goes to bit 3k: built from the binary string, not shifts
    b = format(v, "010b")[::-1]
    return int("".join(c + "00" for c in b)[::-1], 2)


def test_expand_bits_exhaustive(O):
    for v in range(1024):
        assert O.lib().or_expand_bits(v) == _spread_bits_independent(v)


def test_expand_bits_matches_reference(O):
    ref = O.ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    for v in range(1024):
        assert O.lib().or_expand_bits(v) == ref.ref_expand_bits(v)


@pytest.mark.parametrize("xyz,code", [
    ((0.0, 0.0, 0.0), 0),
    ((1.0, 1.0, 1.0), (1 << 30) - 1),         # clamp 1024 -> 1023 on every axis
    ((1.0, 0.0, 0.0), 0x24924924),             # x is the high bit of each triple
    ((0.0, 1.0, 0.0), 0x12492492),
    ((0.0, 0.0, 1.0), 0x09249249),
    ((0.5, 0.0, 0.0), 1 << 29),                # 512 -> bit 9 of x -> bit 29
    ((0.0, 0.5, 0.0), 1 << 28),
    ((0.0, 0.0, 0.5), 1 << 27),
    ((-3.0, 2.0, -0.1), 0x12492492),           # negative clamps to 0, >1 to 1023
    ((1 / 1024, 0.0, 0.0), 4),                 # 1 -> bit 0 of x -> bit 2
    ((float("nan"), 0.0, 0.0), 0),             # undefined in the reference; 0 here
])
def test_morton3d_closed_form(O, xyz, code):
    assert O.lib().or_morton3d(*xyz) == code


def test_morton3d_matches_reference(O):
    ref = O.ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(0)
    pts = rng.uniform(-0.2, 1.2, (20000, 3)).astype(np.float32)
    pts[:64] = np.float32([[0, 1, 0.5]] * 64) + rng.integers(0, 1024, (64, 3)) / 1024
    for x, y, z in pts.tolist():
        assert O.lib().or_morton3d(x, y, z) == ref.ref_morton3d(x, y, z)


def test_ref_affine_codes_match_reference(O, rq):
    """REF_AFFINE: the reference's own (centre + 30) / 1000 -> morton3D
    (accelerators.h:576-580) applied to the oracle's triangle centroids."""
    ref = O.ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    import ctypes as C

    mesh = rq.icosphere(3, 5, center=(3.0, -7.0, -40.0), scale=25.0)
    codes = O.morton_codes(mesh, 0)
    cent = np.empty((mesh.ntri, 3), np.float32)
    (vx, vy, vz, idx), keep = O.mesh_ptrs(mesh)
    out = (C.c_float * 3)()
    for t in range(mesh.ntri):
        O.lib().or_centroid.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_void_p]
        O.lib().or_centroid(vx, vy, vz, idx, t, C.cast(out, C.c_void_p))
        cent[t] = list(out)
    rc = np.empty(mesh.ntri, np.uint32)
    cx, cy, cz = (np.ascontiguousarray(cent[:, k]) for k in range(3))
    ref.ref_morton_centers(O._p(cx), O._p(cy), O._p(cz), mesh.ntri, O._p(rc))
    assert np.array_equal(codes, rc)
    assert len(np.unique(codes)) > 10


def test_scene_aabb_codes_closed_form(O):
    """Point triangles on the corners of a box: centroid bounds are the box,
    so the corner codes are the closed-form extremes."""
    corners = [(a, b, c) for a in (0, 8) for b in (-4, 4) for c in (16, 32)]
    v = []
    for p in corners:
        v += [p, p, p]
    mesh = mesh_from(v, np.arange(len(v)))
    codes = O.morton_codes(mesh, 1)
    expect = []
    for x, y, z in corners:
        qx, qy, qz = (1023 if x == 8 else 0), (1023 if y == 4 else 0), (1023 if z == 32 else 0)
        expect.append(_spread_bits_independent(qx) * 4 + _spread_bits_independent(qy) * 2
                      + _spread_bits_independent(qz))
    assert codes.tolist() == expect


def test_scene_aabb_flat_axis(O):
    """Zero extent on an axis maps that axis to 0 (no division by zero)."""
    v = [(0, 0, 5), (1, 0, 5), (0, 1, 5), (2, 2, 5), (3, 2, 5), (2, 3, 5)]
    mesh = mesh_from(v, np.arange(6))
    codes = O.morton_codes(mesh, 1)
    z_bits = sum(1 << (3 * k) for k in range(10))
    assert all((int(c) & z_bits) == 0 for c in codes)


def _centroids(O, mesh):
    import ctypes as C

    (vx, vy, vz, idx), keep = O.mesh_ptrs(mesh)
    out = (C.c_float * 3)()
    O.lib().or_centroid.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_void_p]
    res = []
    for t in range(mesh.ntri):
        O.lib().or_centroid(vx, vy, vz, idx, t, C.cast(out, C.c_void_p))
        res.append(tuple(np.float32(x) for x in out))
    return res


def test_centroid_closed_forms(O):
    """Triangle centroids with distinct vertices, where the value is fixed:
    exact sums give the exact mean; 1e8 + 1 rounds to 1e8 in float32, so the
    order (v0 + v1) + v2 decides whether the 1 survives; the mean is a
    correctly rounded division by 3 (5 / 3 differs from 5 * float32(1/3))."""
    tris = [((0, 0, 0), (3, 0, 0), (0, 6, 9)),            # (1, 2, 3)
            ((-1.5, 2, 7), (4.5, -8, 1), (0, 0, 1)),      # (1, -2, 3)
            ((1e8, 0, 0), (1, 0, 0), (-1e8, 0, 0)),       # ((1e8 + 1) - 1e8) / 3 = 0
            ((1e8, 0, 0), (-1e8, 0, 0), (1, 0, 0)),       # ((1e8 - 1e8) + 1) / 3 = 1/3
            ((5, 0, 0), (0, 0, 0), (0, 0, 0))]            # 5 / 3 correctly rounded
    v = [p for t in tris for p in t]
    got = _centroids(O, mesh_from(v, np.arange(len(v))))
    third = np.float32(1) / np.float32(3)
    expect = [(1, 2, 3), (1, -2, 3), (0, 0, 0), (third, 0, 0),
              (np.float32(5) / np.float32(3), 0, 0)]
    for g, e in zip(got, expect):
        assert [np.float32(x).view(np.uint32) for x in g] == \
            [np.float32(x).view(np.uint32) for x in e], (g, e)
    assert got[4][0] != np.float32(5) * np.float32(1 / 3)
