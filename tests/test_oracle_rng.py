"""CPU: random-number streams and ray generators of the oracle, pinned by
published known answers, numpy's independent MT19937 and the reference's
own random_double (main.cpp:503-508)."""
import numpy as np
import pytest


def test_philox_published_vectors(O):
    # Random123 known-answer vectors for philox4x32-10
    assert O.philox([0, 0, 0, 0], [0, 0]).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C,
                                                       0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [0x408F276D, 0x41C83B0E,
                                                                     0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                    [0xA4093822, 0x299F31D0]).tolist() == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                           0x24126EA1]


def test_mt19937_standard_known_answer(O):
    # C++ [rand.predef]: the 10000th invocation of a default mt19937 is 4123659995
    assert int(O.mt19937_raw(10000)[-1]) == 4123659995


def test_mt19937_matches_numpy(O):
    raw = O.mt19937_raw(5000)
    rs = np.random.RandomState(5489)  # init_genrand(5489), the default seed
    assert np.array_equal(raw, rs.randint(0, 2 ** 32, size=5000, dtype=np.uint64))


def test_jitter_matches_reference_stream(O):
    ref = O.ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    n = 4096
    mine = O.mt19937_jitter(n)
    theirs = np.empty(n, np.float64)
    ref.ref_random_double(n, O._p(theirs))
    assert np.array_equal(mine, theirs)


def test_camera_rays_geometry(O):
    w, h = 640, 480
    jit = np.full(2 * w * h, 0.5)  # pixel centres
    r = O.camera_rays_ref(w, h, 1, jit)
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert (r["ox"] == 0).all() and (r["tmin"] == 0).all() and np.isinf(r["tmax"]).all()
    ang = np.tan(np.pi * 0.5 * 30 / 180)
    # top-left pixel centre: xx = (2*(0.5/640) - 1)*tan(15)*4/3, yy = (1 - 2*(0.5/480))*tan(15)
    xx, yy = (2 * (0.5 / w) - 1) * ang * w / h, (1 - 2 * (0.5 / h)) * ang
    e = np.array([xx, yy, -1.0]) / np.linalg.norm([xx, yy, -1.0])
    assert np.allclose(d[0], e, atol=1e-6)
    # pixel-centre symmetry about the image centre
    assert np.allclose(d[0, 0], -d[w - 1, 0], atol=1e-7)


def test_uniform_rays(O):
    lo, hi = (-1.0, 2.0, 0.0), (3.0, 4.0, 10.0)
    r = O.uniform_rays(200000, 7, lo, hi)
    o = np.stack([r["ox"], r["oy"], r["oz"]], 1)
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    assert (o >= np.array(lo, np.float32)).all() and (o <= np.array(hi, np.float32)).all()
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.abs(d.mean(axis=0)).max() < 0.01      # isotropic
    assert np.abs((d ** 2).mean(axis=0) - 1 / 3).max() < 0.01
    r2 = O.uniform_rays(1000, 7, lo, hi)
    assert np.array_equal(r2["dx"], r["dx"][:1000])  # counter-based: prefix stable


def test_uniform_rays_at_equals_full_generator(O):
    lo, hi = (-1.0, 2.0, 0.0), (3.0, 4.0, 10.0)
    full = O.uniform_rays(50000, 9, lo, hi)
    ids = np.random.default_rng(1).choice(50000, 3000, replace=False).astype(np.uint32)
    at = O.uniform_rays_at(ids, 9, lo, hi)
    for k in O.RAY_FIELDS:
        assert np.array_equal(at[k].view(np.uint32), full[k][ids].view(np.uint32)), k


def test_uniform_ray_origin_counter_layout(O):
    """Origin x of ray i is lo + ext * (word0 >> 8) / 2^24 of Philox(counter
    (i, 0, 0, 'ORIG'), key seed): recomputed in float32 from the published-
    vector-pinned Philox."""
    lo, hi, seed = (-1.0, 2.0, 0.0), (3.0, 4.0, 10.0), 0x1234567890
    ids = np.array([0, 1, 77, 4095, 1 << 20], np.uint32)
    r = O.uniform_rays_at(ids, seed, lo, hi)
    for k, i in enumerate(ids):
        w = O.philox([int(i), 0, 0, 0x4F524947], [seed & 0xFFFFFFFF, seed >> 32])
        for a, f in enumerate(("ox", "oy", "oz")):
            frac = np.float32(w[a] >> 8) * np.float32(1.0 / 16777216.0)
            e = np.float32(lo[a]) + np.float32(np.float32(hi[a]) - np.float32(lo[a])) * frac
            assert r[f][k] == e, (i, f)


def _pinhole_pixels(r, eye, look, up, fov, w, h):
    """Image coordinates (px, py) at which each ray crosses the image plane,
    from an independent float64 camera basis."""
    f = np.asarray(look, np.float64) - np.asarray(eye, np.float64)
    f /= np.linalg.norm(f)
    rt = np.cross(f, np.asarray(up, np.float64))
    rt /= np.linalg.norm(rt)
    u = np.cross(rt, f)
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    th = np.tan(np.radians(fov) / 2)
    sx = (d @ rt) / (d @ f)
    sy = (d @ u) / (d @ f)
    px = (sx / (th * w / h) + 1) / 2 * w
    py = (1 - sy / th) / 2 * h
    return px, py


def test_pinhole_rays_pixel_footprint_and_jitter(O):
    """Ray i crosses the image plane inside pixel (i % w, i / w) of a w x
    ceil(n/w) image (row-major, y down, right = f x up), at the jitter given
    by the first two Philox words of counter (i, 0, 0, 'PINH')."""
    eye, look, up, fov, w = (1.0, 2.0, 3.0), (1.5, 1.0, -4.0), (0.0, 1.0, 0.0), 40.0, 64
    n = 64 * 48 - 5                     # ragged last row
    h = -(-n // w)
    seed = 77
    r = O.pinhole_rays_at(None, n, seed, eye, look, up, fov, w)
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert (r["ox"] == np.float32(eye[0])).all() and (r["oz"] == np.float32(eye[2])).all()
    assert (r["tmin"] == 0).all() and np.isinf(r["tmax"]).all()
    px, py = _pinhole_pixels(r, eye, look, up, fov, w, h)
    i = np.arange(n)
    jx, jy = px - i % w, py - i // w
    assert (jx > -1e-3).all() and (jx < 1 + 1e-3).all()
    assert (jy > -1e-3).all() and (jy < 1 + 1e-3).all()
    for k in (0, 1, 63, 64, 1000, n - 1):
        wd = O.philox([k, 0, 0, 0x50494E48], [seed, 0])
        assert abs(jx[k] - (wd[0] >> 8) / 2.0 ** 24) < 2e-3, k
        assert abs(jy[k] - (wd[1] >> 8) / 2.0 ** 24) < 2e-3, k


def test_pinhole_rays_at_equals_full_generator(O):
    args = (5000, 3, (0.0, 0.0, 4.0), (0.1, -0.2, 0.0), (0.0, 1.0, 0.0), 35.0, 100)
    full = O.pinhole_rays_at(None, *args)
    ids = np.array([0, 99, 100, 2500, 4999], np.uint32)
    at = O.pinhole_rays_at(ids, *args)
    for k in O.RAY_FIELDS:
        assert np.array_equal(at[k].view(np.uint32), full[k][ids].view(np.uint32)), k
