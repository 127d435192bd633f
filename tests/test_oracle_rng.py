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
