"""CPU: the host check that rq_bvh_load makes before any upload
(rq_bvh_blob_validate). Blobs are written here from the oracle's tree in the
rq_bvh_save format (helpers.blob_payload), so the check is exercised without
a device: a well-formed blob is accepted, and every crafted corruption, with
its checksum recomputed so that only the payload check can catch it, is
refused with RQ_E_IO; a blob of other primitives is RQ_E_INVALID."""
import numpy as np
import pytest

from helpers import blob_from_payload, blob_payload, duplicate_mesh, random_soup

NODE, TRI = 64, 48


def _blob(O, rq, prims):
    ob = O.OracleBVH(prims, rq.MORTON_SCENE_AABB)
    arr = ob.arrays()
    return blob_payload(prims, arr, ob.height), arr, ob.height


def _n(prims):
    return prims.n if hasattr(prims, "cx") else prims.ntri


class Layout:
    """Byte offsets of the payload arrays."""

    def __init__(self, n):
        self.codes = 0
        self.sidx = 4 * n
        self.parent = 8 * n
        self.nodes = self.parent + 4 * (2 * n - 1)
        self.tris = self.nodes + NODE * (n - 1)
        self.root = self.tris + TRI * n
        self.height = self.root + 24


def _i32(buf, off, val=None):
    a = np.frombuffer(buf, np.int32, 1, off) if val is None else None
    if val is None:
        return int(a[0])
    buf[off:off + 4] = np.int32(val).tobytes()


def _f32(buf, off, val):
    buf[off:off + 4] = np.float32(val).tobytes()


def _spheres(rq, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    return rq.SphereArrays(c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy(),
                           rng.uniform(0.02, 0.2, n).astype(np.float32),
                           np.zeros(3 * n, np.float32))


@pytest.mark.parametrize("which", ["ico2", "dup64", "soup500", "tri1", "tri2", "spheres"])
def test_wellformed_blob_accepted(O, rq, which):
    prims = {"ico2": lambda: rq.icosphere(2, 1, center=(0, 0, -5)),
             "dup64": lambda: duplicate_mesh(64),
             "soup500": lambda: random_soup(500, 3, spread=1.5, size=0.05),
             "tri1": lambda: random_soup(1, 5),
             "tri2": lambda: random_soup(2, 6),
             "spheres": lambda: _spheres(rq, 300, 4)}[which]()
    payload, _, _ = _blob(O, rq, prims)
    rc, msg = rq.bvh_blob_validate(blob_from_payload(prims, payload), prims)
    assert rc == rq.RQ_OK, msg


def _corruptions(arr, n, height):
    """(name, function editing the payload in place)."""
    L = Layout(n)
    left, right = arr["left"], arr["right"]
    # an internal node with an internal left child and one with a leaf child
    i_int = int(np.nonzero(left >= 0)[0][0])
    i_leaf = int(np.nonzero(right < 0)[0][0])
    j_leaf = int(np.nonzero(right < 0)[0][1])
    lnk = lambda i, side: L.nodes + NODE * i + (12 if side == 0 else 28)  # noqa: E731
    return [
        ("link past the nodes", lambda b: _i32(b, lnk(i_int, 0), n - 1)),
        ("link to the root (a cycle)", lambda b: _i32(b, lnk(i_int, 0), 0)),
        ("leaf link past the leaves", lambda b: _i32(b, lnk(i_leaf, 1), ~n)),
        ("leaf reached twice", lambda b: _i32(b, lnk(i_leaf, 1), int(right[j_leaf]))),
        ("parent link", lambda b: _i32(b, L.parent + 4 * 5, _i32(b, L.parent + 4 * 5) ^ 1)),
        ("sorted order not a permutation", lambda b: _i32(b, L.sidx + 4, _i32(b, L.sidx))),
        ("primitive id past n", lambda b: _i32(b, L.tris + TRI * 3 + 12, n + 7)),
        ("primitive id of another leaf", lambda b: _i32(b, L.tris + 12, _i32(b, L.tris + TRI + 12))),
        ("leaf vertex moved", lambda b: _f32(b, L.tris + TRI * 2,
                                             np.frombuffer(b, np.float32, 1, L.tris + TRI * 2)[0] + 0.5)),
        ("child box shrunk", lambda b: _f32(b, L.nodes + NODE * i_int + 4,
                                            np.frombuffer(b, np.float32, 1,
                                                          L.nodes + NODE * i_int + 4)[0] + 0.25)),
        ("root box", lambda b: _f32(b, L.root, np.frombuffer(b, np.float32, 1, L.root)[0] - 1)),
        ("height off by one", lambda b: _i32(b, L.height, height + 1)),
        ("height beyond the stack", lambda b: _i32(b, L.height, 64)),
    ]


@pytest.mark.parametrize("which", ["ico2", "spheres"])
def test_crafted_blobs_refused(O, rq, which):
    prims = rq.icosphere(2, 1, center=(0, 0, -5)) if which == "ico2" else _spheres(rq, 300, 4)
    payload, arr, height = _blob(O, rq, prims)
    n = _n(prims)
    for name, edit in _corruptions(arr, n, height):
        bad = bytearray(payload)
        edit(bad)
        assert bad != payload, name
        rc, msg = rq.bvh_blob_validate(blob_from_payload(prims, bad), prims)
        assert rc == rq.RQ_E_IO, f"{name}: accepted ({rc}, {msg})"
        assert msg, name


def test_damaged_and_foreign_blobs(O, rq):
    prims = rq.icosphere(2, 1, center=(0, 0, -5))
    payload, _, _ = _blob(O, rq, prims)
    good = blob_from_payload(prims, payload)
    assert rq.bvh_blob_validate(good[:40], prims)[0] == rq.RQ_E_IO           # truncated header
    assert rq.bvh_blob_validate(good[:-1], prims)[0] == rq.RQ_E_IO           # truncated payload
    assert rq.bvh_blob_validate(b"X" + good[1:], prims)[0] == rq.RQ_E_IO     # magic
    flipped = bytearray(good)
    flipped[100] ^= 1                                                          # checksum not redone
    assert rq.bvh_blob_validate(bytes(flipped), prims)[0] == rq.RQ_E_IO
    other = rq.icosphere(2, 2, center=(0, 0, -5))                             # same n, other vertices
    assert rq.bvh_blob_validate(good, other)[0] == rq.RQ_E_INVALID
    assert rq.bvh_blob_validate(good, rq.icosphere(3, 1))[0] == rq.RQ_E_INVALID  # other n
