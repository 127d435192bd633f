"""The five benchmark workloads as seeded, deterministic inputs, shared by
bench.py (which times them) and the full-size parity tests (which check them).

c1  plain method: displaced icosphere L6 (81,920 triangles) at (0,0,-5),
    640x480, 1 spp with the reference's mt19937 jitter (main.cpp:503-508).
c2  incoherent large mesh: icosphere L8 (1,310,720 triangles) at (0,0,-2),
    1920x1080x4 spp Philox jitter, one mirror bounce and shadow rays.
c3  raw ray batch: 256 L5 icospheres (5,242,880 triangles) with Poisson-disk
    centres in [0,100]^3; 2^24 uniform rays in the scene box inflated 10%.
c4  512 meshes, F log-uniform in [1e3, 5e5] (icospheres and soups), R
    log-uniform in [2^16, 2^20] uniform rays per mesh.
c5  64 resident meshes, F log-uniform in [1e4, 2e6]; requests alternate
    pinhole rays aimed at the mesh and uniform rays in its inflated box.
"""
from __future__ import annotations

import numpy as np

from . import (CAMERA_PHILOX, icosphere, icosphere_field, render_defaults, soup)

C3_RAYS = 1 << 24


def inflated_box(mesh, frac: float = 0.1) -> tuple[np.ndarray, np.ndarray]:
    """The mesh's bounding box grown by `frac` of its extent on every side."""
    lo, hi = mesh.bounds()
    ext = hi - lo
    return ((lo - ext * np.float32(frac)).astype(np.float32),
            (hi + ext * np.float32(frac)).astype(np.float32))


def c1():
    return icosphere(6, 1, center=(0.0, 0.0, -5.0)), render_defaults(), "c1_plain_640x480"


def c2():
    mesh = icosphere(8, 2, center=(0.0, 0.0, -2.0))
    p = render_defaults()
    p.width, p.height, p.spp = 1920, 1080, 4
    p.camera, p.seed, p.bounce = CAMERA_PHILOX, 2, 1
    return mesh, p, "c2_incoherent_large_mesh"


def c3():
    """Mesh, ray box (lo, hi), ray count and ray seed of the raw ray batch."""
    mesh = icosphere_field(256, 5, 3, 3, 100.0, 2.5)
    lo, hi = inflated_box(mesh)
    return mesh, (lo, hi), C3_RAYS, 3


def _level_at_most(f: int) -> int:
    """The largest icosphere level L with 20 * 4^L <= f (0 if none)."""
    level = 0
    while 20 * 4 ** (level + 1) <= f:
        level += 1
    return level


def c4_sizes(count: int = 512) -> tuple[np.ndarray, np.ndarray]:
    """Triangle targets F_m = round(exp(U[ln 1e3, ln 5e5])) (numpy seed 4) and
    ray counts R_m = round(exp(U[ln 2^16, ln 2^20])) (numpy seed 5)."""
    tris = np.rint(np.exp(np.random.default_rng(4).uniform(np.log(1e3), np.log(5e5), count)))
    nrays = np.rint(np.exp(np.random.default_rng(5).uniform(np.log(65536.0), np.log(1048576.0),
                                                             count)))
    return tris.astype(np.int64), nrays.astype(np.int64)


def c4_mesh(m: int, f: int):
    """Mesh m of the batch: even m a displaced icosphere at the largest level
    with 20*4^L <= F_m (seed 1000 + m), odd m a soup of F_m triangles
    (centroids uniform in the unit cube, edge ~N(0.02, 0.005) clipped at
    0.002; seed 2000 + m)."""
    if m % 2 == 0:
        return icosphere(_level_at_most(int(f)), 1000 + m)
    return soup(int(f), 2000 + m)


def c4_ray_seed(m: int) -> int:
    return 3000 + m


def c5_meshes(count: int, seed: int, lo: float = 1e4, hi: float = 2e6):
    """F ~ round(exp(U[ln lo, ln hi])) (numpy seed); even m a displaced
    icosphere at the largest level with 20*4^L <= F, odd m a random soup."""
    rng = np.random.default_rng(seed)
    tris = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    out = []
    for m in range(count):
        if m % 2 == 0:
            out.append(icosphere(_level_at_most(int(tris[m])), 10 * seed + m))
        else:
            out.append(soup(int(tris[m]), 1000 * seed + m))
    return out


def c5_rays_spec(bounds, kind: int, seed: int) -> dict:
    """Ray parameters of a serving request on a mesh with bounds (lo, hi):
    kind 0 pinhole rays from 2.5 box radii in front of the centre looking at
    it (fov 40, 1024 wide), kind 1 uniform rays in the box inflated 10%."""
    lo, hi = bounds
    c = (lo + hi) * 0.5
    r = float(np.linalg.norm(hi - lo)) * 0.5
    if kind == 0:
        return dict(pinhole=dict(eye=(float(c[0]), float(c[1]), float(c[2] + 2.5 * r)),
                                 look_at=tuple(float(x) for x in c), up=(0.0, 1.0, 0.0),
                                 fov_deg=40.0, width=1024), seed=seed)
    ext = hi - lo
    return dict(box=(tuple(float(x) for x in lo - 0.1 * ext),
                     tuple(float(x) for x in hi + 0.1 * ext)), seed=seed)
