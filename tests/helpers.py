"""Shared inputs for the parity tests (seeded, deterministic)."""
from __future__ import annotations

import numpy as np

import rq

RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def mesh_from(v: np.ndarray, f: np.ndarray) -> rq.MeshArrays:
    v = np.asarray(v, np.float32).reshape(-1, 3)
    return rq.MeshArrays(v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy(),
                         np.asarray(f, np.uint32).reshape(-1).copy())


def random_soup(n: int, seed: int, spread: float = 1.0, size: float = 0.05) -> rq.MeshArrays:
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    v = c + rng.normal(0, size, (n, 3, 3))
    return mesh_from(v.reshape(-1, 3), np.arange(3 * n))


def duplicate_mesh(n: int) -> rq.MeshArrays:
    """n identical triangles: every Morton code equal, every hit a tie."""
    v = np.tile(np.array([[-1, -1, -5], [1, -1, -5], [0, 1, -5]], np.float32), (n, 1))
    return mesh_from(v, np.arange(3 * n))


def rays_toward(n: int, seed: int, target=(0.0, 0.0, -5.0), spread=1.5, origin=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    o = np.tile(np.asarray(origin, np.float32), (n, 1))
    tgt = np.asarray(target, np.float32) + rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    return dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, np.float32),
                tmax=np.full(n, np.inf, np.float32))


def random_rays(n: int, seed: int, lo=(-2, -2, -7), hi=(2, 2, -3)):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    return dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, np.float32),
                tmax=np.full(n, np.inf, np.float32))


def concat_rays(*rs):
    return {k: np.concatenate([r[k] for r in rs]).astype(np.float32) for k in RAY_FIELDS}


def edge_rays():
    """Axis-aligned directions (zero components), negative zeros, clipped and
    inactive intervals."""
    rows = []
    for d in [(0, 0, -1), (0, 0, 1), (1, 0, 0), (0, -1, 0), (-0.0, -0.0, -1), (0, 0.6, -0.8)]:
        for o in [(0, 0, 0), (0.1, -0.2, 0), (0, 0, -5), (0.5, 0.5, -2)]:
            for tmin, tmax in [(0, np.inf), (0, 4.5), (4.0, 6.0), (3, 3), (5, 1)]:
                rows.append((*o, *d, tmin, tmax))
    a = np.array(rows, np.float32)
    return {k: a[:, i].copy() for i, k in enumerate(RAY_FIELDS)}


def nonfinite_rays():
    """Rays with a NaN or infinite origin or direction component (one axis,
    or all of them), each after a copy of itself that is finite and aimed at
    the origin region; none of the non-finite rays can have a finite hit."""
    rows = []
    bad = (np.nan, np.inf, -np.inf)
    base_o, base_d = (0.1, -0.2, 0.0), (0.0, 0.0, -1.0)
    for v in bad:
        for which in ("o", "d"):
            for axes in ((0,), (1,), (2,), (0, 1, 2)):
                o, d = list(base_o), list(base_d)
                for a in axes:
                    (o if which == "o" else d)[a] = v
                rows.append((*base_o, *base_d, 0.0, np.inf))
                rows.append((*o, *d, 0.0, np.inf))
    a = np.array(rows, np.float32)
    return {k: a[:, i].copy() for i, k in enumerate(RAY_FIELDS)}


def mesh_configs():
    """Name -> mesh builder for the small parity sweep."""
    return {
        "tri1": lambda: mesh_from([[-1, -1, -5], [1, -1, -5], [0, 1, -5]], [0, 1, 2]),
        "tri2": lambda: mesh_from([[-1, -1, -5], [1, -1, -5], [0, 1, -5], [-1, -1, -6],
                                   [1, -1, -6], [0, 1, -6]], np.arange(6)),
        "dup64": lambda: duplicate_mesh(64),
        "soup100": lambda: random_soup(100, 1, spread=1.5),
        "ico2": lambda: rq.icosphere(2, 1, center=(0, 0, -5)),
        "soup4097": lambda: random_soup(4097, 3, spread=1.5, size=0.02),
        "ico5": lambda: rq.icosphere(5, 2, center=(0, 0, -5)),
    }


# ---- BVH blob written from the oracle's tree (the rq_bvh_save format) ----
BLOB_MAGIC = b"RQBVHBL1"


def fnv1a64(data: bytes, h: int = 0xCBF29CE484222325) -> int:
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def _prims_fingerprint(prims) -> int:
    h = 0xCBF29CE484222325
    if hasattr(prims, "cx"):
        for a in (prims.cx, prims.cy, prims.cz, prims.r):
            h = fnv1a64(np.ascontiguousarray(a, np.float32).tobytes(), h)
        return h
    for a in (prims.vx, prims.vy, prims.vz):
        h = fnv1a64(np.ascontiguousarray(a, np.float32).tobytes(), h)
    return fnv1a64(np.ascontiguousarray(prims.idx, np.uint32).tobytes(), h)


def blob_payload(prims, arrays: dict, height: int) -> bytearray:
    """The blob payload of a tree given as the oracle's arrays (codes,
    sorted_idx, left, right, boxes[2n-1, 6]: internal nodes, then leaves)."""
    codes, sidx = arrays["codes"], arrays["sorted_idx"]
    left, right, boxes = arrays["left"], arrays["right"], arrays["boxes"].reshape(-1, 6)
    n = sidx.size
    ni = n - 1
    parent = np.full(2 * n - 1, -1, np.int32)
    nodes = np.zeros((ni, 16), np.float32)
    nv = nodes.view(np.int32)
    for i in range(ni):
        for side, c in enumerate((int(left[i]), int(right[i]))):
            u = c if c >= 0 else ni + ~c
            parent[u] = 2 * i + side
            nodes[i, 8 * side:8 * side + 3] = boxes[u, :3]
            nodes[i, 8 * side + 4:8 * side + 7] = boxes[u, 3:]
        nv[i, 3], nv[i, 7] = left[i], right[i]
    tris = np.zeros((n, 12), np.float32)
    tv = tris.view(np.int32)
    p = sidx.astype(np.int64)
    if hasattr(prims, "cx"):
        r = np.asarray(prims.r, np.float32)[p]
        tris[:, 0], tris[:, 1], tris[:, 2] = (np.asarray(a, np.float32)[p]
                                              for a in (prims.cx, prims.cy, prims.cz))
        tris[:, 4], tris[:, 5] = r, r * r
    else:
        f = np.asarray(prims.idx, np.int64).reshape(-1, 3)[p]
        v = np.stack([np.asarray(a, np.float32) for a in (prims.vx, prims.vy, prims.vz)], 1)
        v0, v1, v2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
        tris[:, 0:3], tris[:, 4:7], tris[:, 8:11] = v0, v1 - v0, v2 - v0
    tv[:, 3] = sidx.astype(np.int32)
    out = bytearray()
    for a in (codes.astype(np.uint32), sidx.astype(np.uint32), parent, nodes, tris,
              boxes[0].astype(np.float32), np.array([height], np.uint32)):
        out += np.ascontiguousarray(a).tobytes()
    return out


def blob_from_payload(prims, payload: bytes, mode: int = 1) -> bytes:
    n = prims.n if hasattr(prims, "cx") else prims.ntri
    kind = 1 if hasattr(prims, "cx") else 0
    elems = n if kind else prims.nv
    hdr = (BLOB_MAGIC + np.array([1, kind, mode, n, elems, 0], np.uint32).tobytes() +
           np.array([_prims_fingerprint(prims), len(payload), fnv1a64(bytes(payload)), 0],
                    np.uint64).tobytes())
    assert len(hdr) == 64
    return hdr + bytes(payload)
