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
