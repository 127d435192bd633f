"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference build
(_ref/libref.so). TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke
check and the CPU-baseline leg of bench.py; never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF = HERE / "_ref" / "libref.so"


def build() -> None:
    """Compile the oracle (and the reference, when its sources are present)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "all"], check=True)


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("prim", "t", "u", "v")]


class ShadeParams(C.Structure):
    _fields_ = [("light", C.c_float * 3), ("emission", C.c_float * 3),
                ("surface", C.c_float * 3), ("sky", C.c_float * 3), ("bias", C.c_float),
                ("bounce", C.c_int), ("shadows", C.c_int), ("prim_rgb", C.c_void_p),
                ("prim_mat", C.c_void_p), ("mode", C.c_int)]


RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")
_lib = None
_ref = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp, u32, u64, i32, f = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float
        sig = {
            "or_expand_bits": (u32, [u32]),
            "or_morton3d": (u32, [f, f, f]),
            "or_morton_codes": (i32, [vp, vp, vp, u32, vp, u32, i32, vp]),
            "or_sort_pairs": (None, [vp, u32, vp, vp]),
            "or_karras": (None, [vp, u32, vp, vp]),
            "or_topdown": (None, [vp, u32, vp, vp]),
            "or_refit": (None, [vp, vp, vp, vp, vp, u32, vp, vp, vp]),
            "or_tree_height": (u32, [vp, vp, u32]),
            "or_mt": (i32, [vp, vp, vp, vp, vp, f, f, vp, vp, vp]),
            "or_ray_sphere": (i32, [vp, f, vp, vp, vp, vp]),
            "or_closest_brute": (None, [vp, vp, vp, vp, u32, C.POINTER(Rays), u32,
                                        C.POINTER(Hits), i32]),
            "or_any_brute": (None, [vp, vp, vp, vp, u32, C.POINTER(Rays), u32, vp, i32]),
            "or_bvh_build": (vp, [vp, vp, vp, u32, vp, u32, i32]),
            "or_bvh_build_spheres": (vp, [vp, vp, vp, vp, u32, i32]),
            "or_closest_brute_spheres": (None, [vp, vp, vp, vp, u32, C.POINTER(Rays), u32,
                                                C.POINTER(Hits), i32, i32]),
            "or_any_brute_spheres": (None, [vp, vp, vp, vp, u32, C.POINTER(Rays), u32, vp, i32]),
            "or_bvh_free": (None, [vp]),
            "or_bvh_arrays": (None, [vp, vp, vp, vp, vp, vp]),
            "or_bvh_height": (u32, [vp]),
            "or_bvh_treelets": (u32, [vp]),
            "or_treelet_dp": (f, [vp, i32, vp, vp]),
            "or_closest_bvh": (None, [vp, C.POINTER(Rays), u32, C.POINTER(Hits), vp, vp, i32]),
            "or_any_bvh": (None, [vp, C.POINTER(Rays), u32, vp, vp, vp, i32]),
            "or_closest_bvh_stack": (None, [vp, C.POINTER(Rays), u32, vp, i32]),
            "or_mt19937_jitter": (None, [u32, vp]),
            "or_mt19937_raw": (None, [u32, vp]),
            "or_philox4x32_10": (None, [vp, vp, vp]),
            "or_camera_rays_ref": (None, [u32, u32, u32, vp] + [vp] * 8),
            "or_camera_rays_philox": (None, [u32, u32, u32, u64] + [vp] * 8),
            "or_uniform_rays": (None, [u32, u64, vp, vp] + [vp] * 8),
            "or_uniform_rays_at": (None, [vp, u32, u64, vp, vp] + [vp] * 8),
            "or_pinhole_rays_at": (None, [vp, u32, u32, u64, vp, vp, vp, f, u32] + [vp] * 8),
            "or_render_samples": (None, [vp, vp, vp, vp, vp, C.POINTER(Rays), u32,
                                         C.POINTER(ShadeParams), vp, vp, vp, vp, vp, vp, i32]),
            "or_resolve_ppm": (None, [vp, u32, u32, vp]),
            "port_bvh_build": (vp, [vp, vp, vp, vp, u32, vp]),
            "port_lbvh_build": (vp, [vp, vp, vp, vp, u32, vp]),
            "port_bvh_free": (None, [vp]),
            "port_closest": (None, [vp] + [vp] * 8 + [u32, vp, vp, i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref() -> C.CDLL | None:
    """The reference's own code (built from its sources by oracle/Makefile), or None."""
    global _ref
    if _ref is None and REF.exists():
        L = C.CDLL(str(REF))
        vp, u32, i32, f = C.c_void_p, C.c_uint, C.c_int, C.c_float
        for name, res, args in [
            ("ref_expand_bits", u32, [u32]),
            ("ref_morton3d", u32, [f, f, f]),
            ("ref_morton_centers", None, [vp, vp, vp, u32, vp]),
            ("ref_ray_sphere", i32, [vp, f, vp, vp, vp, vp]),
            ("ref_box_hit", i32, [vp, vp, vp, vp]),
            ("ref_sphere_closest", i32, [vp] * 4 + [u32] + [vp] * 6 + [u32, i32, vp, vp]),
            ("ref_random_double", None, [u32, vp]),
            ("ref_sphere_scene_colors", None, [vp] * 5 + [u32] + [vp] * 6 + [u32, vp]),
            ("ref_sphere_scene_colors_mat", None, [vp] * 6 + [u32] + [vp] * 6 + [u32, vp]),
            ("ref_kdtree_any", None, [vp] * 4 + [u32] + [vp] * 6 + [u32, vp]),
            ("ref_camera_dirs", None, [u32, u32, u32, vp, vp, vp, vp]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _ref = L
    return _ref


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def rays_struct(rays: dict) -> tuple[Rays, list]:
    keep = [_f32(rays[k]) for k in RAY_FIELDS]
    return Rays(*(_p(a) for a in keep)), keep


def mesh_ptrs(mesh) -> tuple[list, list]:
    keep = [_f32(mesh.vx), _f32(mesh.vy), _f32(mesh.vz),
            np.ascontiguousarray(mesh.idx, np.uint32)]
    return [_p(a) for a in keep], keep


# ------------------------------------------------------------------ wrappers
def morton_codes(mesh, mode: int) -> np.ndarray:
    (vx, vy, vz, idx), keep = mesh_ptrs(mesh)
    out = np.empty(mesh.ntri, np.uint32)
    lib().or_morton_codes(vx, vy, vz, mesh.nv, idx, mesh.ntri, mode, _p(out))
    return out


def sort_pairs(codes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    codes = np.ascontiguousarray(codes, np.uint32)
    sc = np.empty_like(codes)
    si = np.empty_like(codes)
    lib().or_sort_pairs(_p(codes), codes.size, _p(sc), _p(si))
    return sc, si


def karras(sorted_codes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    sc = np.ascontiguousarray(sorted_codes, np.uint32)
    n = sc.size
    l = np.zeros(max(n - 1, 0), np.int32)
    r = np.zeros(max(n - 1, 0), np.int32)
    lib().or_karras(_p(sc), n, _p(l), _p(r))
    return l, r


def topdown(sorted_codes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    sc = np.ascontiguousarray(sorted_codes, np.uint32)
    n = sc.size
    l = np.zeros(max(n - 1, 0), np.int32)
    r = np.zeros(max(n - 1, 0), np.int32)
    lib().or_topdown(_p(sc), n, _p(l), _p(r))
    return l, r


def mt(v0, v1, v2, o, d, tmin=0.0, tmax=np.inf):
    a = [_f32(x) for x in (v0, v1, v2, o, d)]
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = lib().or_mt(*(_p(x) for x in a), tmin, tmax, C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def sphere_ptrs(sph) -> tuple[list, list]:
    keep = [_f32(sph.cx), _f32(sph.cy), _f32(sph.cz), _f32(sph.r)]
    return [_p(a) for a in keep], keep


def treelet_dp(boxes: np.ndarray) -> tuple[float, np.ndarray, np.ndarray]:
    """Optimal tree over m <= 7 boxes (m x 6): cost, split and height per subset."""
    b = np.ascontiguousarray(boxes, np.float32)
    m = len(b)
    split = np.zeros(1 << m, np.uint8)
    height = np.zeros(1 << m, np.uint8)
    c = lib().or_treelet_dp(_p(b), m, _p(split), _p(height))
    return float(np.float32(c)), split, height


class OracleBVH:
    """The oracle's LBVH: Morton -> stable sort -> Karras -> refit. `mesh` is a
    triangle mesh (vx, vy, vz, idx) or a sphere set (cx, cy, cz, r)."""

    def __init__(self, mesh, mode: int = 1):
        if hasattr(mesh, "cx"):
            (cx, cy, cz, r), self._keep = sphere_ptrs(mesh)
            self.n = len(self._keep[0])
            self.h = lib().or_bvh_build_spheres(cx, cy, cz, r, self.n, mode)
            return
        (vx, vy, vz, idx), self._keep = mesh_ptrs(mesh)
        self.n = mesh.ntri
        self.h = lib().or_bvh_build(vx, vy, vz, mesh.nv, idx, mesh.ntri, mode)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_bvh_free(self.h)
            self.h = None

    def arrays(self) -> dict:
        n = self.n
        out = dict(codes=np.empty(n, np.uint32), sorted_idx=np.empty(n, np.uint32),
                   left=np.empty(max(n - 1, 0), np.int32),
                   right=np.empty(max(n - 1, 0), np.int32),
                   boxes=np.empty(max(2 * n - 1, 0) * 6, np.float32))
        lib().or_bvh_arrays(self.h, *(_p(out[k]) for k in
                                      ("codes", "sorted_idx", "left", "right", "boxes")))
        out["boxes"] = out["boxes"].reshape(-1, 6)
        return out

    @property
    def height(self) -> int:
        return int(lib().or_bvh_height(self.h))

    def treelets(self) -> int:
        """Treelet restructuring in place (or_bvh_treelets); treelets replaced."""
        return int(lib().or_bvh_treelets(self.h))

    def closest(self, rays: dict, threads: int = 8) -> dict:
        rs, keep = rays_struct(rays)
        n = len(keep[0])
        out = dict(prim=np.empty(n, np.int32), t=np.empty(n, np.float32),
                   u=np.empty(n, np.float32), v=np.empty(n, np.float32),
                   visits=np.empty(n, np.uint32), tests=np.empty(n, np.uint32))
        h = Hits(_p(out["prim"]), _p(out["t"]), _p(out["u"]), _p(out["v"]))
        lib().or_closest_bvh(self.h, C.byref(rs), n, C.byref(h), _p(out["visits"]),
                             _p(out["tests"]), threads)
        return out

    def max_stack(self, rays: dict, threads: int = 8) -> np.ndarray:
        """Deepest stack of each ray's canonical closest-hit walk."""
        rs, keep = rays_struct(rays)
        n = len(keep[0])
        out = np.empty(n, np.uint32)
        lib().or_closest_bvh_stack(self.h, C.byref(rs), n, _p(out), threads)
        return out

    def any(self, rays: dict, threads: int = 8) -> dict:
        rs, keep = rays_struct(rays)
        n = len(keep[0])
        out = dict(occ=np.empty(n, np.uint8), visits=np.empty(n, np.uint32),
                   tests=np.empty(n, np.uint32))
        lib().or_any_bvh(self.h, C.byref(rs), n, _p(out["occ"]), _p(out["visits"]),
                         _p(out["tests"]), threads)
        return out


def closest_brute(mesh, rays: dict, threads: int = 8, box_filter: bool = False) -> dict:
    """NONE loop. Spheres: box_filter keeps only spheres whose box the ray hits."""
    rs, rk = rays_struct(rays)
    n = len(rk[0])
    out = dict(prim=np.empty(n, np.int32), t=np.empty(n, np.float32),
               u=np.empty(n, np.float32), v=np.empty(n, np.float32))
    h = Hits(_p(out["prim"]), _p(out["t"]), _p(out["u"]), _p(out["v"]))
    if hasattr(mesh, "cx"):
        (cx, cy, cz, r), keep = sphere_ptrs(mesh)
        lib().or_closest_brute_spheres(cx, cy, cz, r, len(keep[0]), C.byref(rs), n, C.byref(h),
                                       1 if box_filter else 0, threads)
        return out
    (vx, vy, vz, idx), keep = mesh_ptrs(mesh)
    lib().or_closest_brute(vx, vy, vz, idx, mesh.ntri, C.byref(rs), n, C.byref(h), threads)
    return out


def any_brute(mesh, rays: dict, threads: int = 8) -> np.ndarray:
    rs, rk = rays_struct(rays)
    n = len(rk[0])
    occ = np.empty(n, np.uint8)
    if hasattr(mesh, "cx"):
        (cx, cy, cz, r), keep = sphere_ptrs(mesh)
        lib().or_any_brute_spheres(cx, cy, cz, r, len(keep[0]), C.byref(rs), n, _p(occ), threads)
        return occ
    (vx, vy, vz, idx), keep = mesh_ptrs(mesh)
    lib().or_any_brute(vx, vy, vz, idx, mesh.ntri, C.byref(rs), n, _p(occ), threads)
    return occ


def _ray_out(n: int) -> dict:
    return {k: np.empty(n, np.float32) for k in RAY_FIELDS}


def camera_rays_ref(w: int, h: int, spp: int, jitter: np.ndarray) -> dict:
    out = _ray_out(w * h * spp)
    j = np.ascontiguousarray(jitter, np.float64)
    lib().or_camera_rays_ref(w, h, spp, _p(j), *(_p(out[k]) for k in RAY_FIELDS))
    return out


def camera_rays_philox(w: int, h: int, spp: int, seed: int) -> dict:
    out = _ray_out(w * h * spp)
    lib().or_camera_rays_philox(w, h, spp, seed, *(_p(out[k]) for k in RAY_FIELDS))
    return out


def uniform_rays(n: int, seed: int, lo, hi) -> dict:
    out = _ray_out(n)
    lo3, hi3 = _f32(lo), _f32(hi)
    lib().or_uniform_rays(n, seed, _p(lo3), _p(hi3), *(_p(out[k]) for k in RAY_FIELDS))
    return out


def uniform_rays_at(ids, seed: int, lo, hi) -> dict:
    """Rays ids[k] of uniform_rays(., seed, lo, hi) (any n above max(ids))."""
    ids = np.ascontiguousarray(ids, np.uint32)
    out = _ray_out(ids.size)
    lo3, hi3 = _f32(lo), _f32(hi)
    lib().or_uniform_rays_at(_p(ids), ids.size, seed, _p(lo3), _p(hi3),
                             *(_p(out[k]) for k in RAY_FIELDS))
    return out


def pinhole_rays_at(ids, n: int, seed: int, eye, look_at, up=(0.0, 1.0, 0.0),
                    fov_deg: float = 40.0, width: int = 1024) -> dict:
    """Rays ids[k] of the serving pinhole camera for a request of n rays
    (ids None: all n)."""
    if ids is None:
        ids_p, count = None, n
    else:
        ids = np.ascontiguousarray(ids, np.uint32)
        ids_p, count = _p(ids), ids.size
    out = _ray_out(count)
    e, la, u = _f32(eye), _f32(look_at), _f32(up)
    lib().or_pinhole_rays_at(ids_p, count, n, seed, _p(e), _p(la), _p(u), fov_deg, width,
                             *(_p(out[k]) for k in RAY_FIELDS))
    return out


def mt19937_jitter(n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().or_mt19937_jitter(n, _p(out))
    return out


def mt19937_raw(n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().or_mt19937_raw(n, _p(out))
    return out


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def shade_params(light=(0, 3, 30), emission=(1, 1, 1), surface=(0, 0, 0), sky=(0.6, 0.8, 1.0),
                 bias=1e-4, bounce=0, shadows=1, prim_rgb=None, prim_mat=None,
                 mode=0) -> ShadeParams:
    """prim_rgb: float32 array (3 per primitive), prim_mat: uint8 material per
    primitive (0 diffuse, 1 reflection + refraction, 2 reflection); both kept
    alive by the returned object. mode 1: castRay's KDTREE case (black on any hit)."""
    sp = ShadeParams((C.c_float * 3)(*light), (C.c_float * 3)(*emission),
                     (C.c_float * 3)(*surface), (C.c_float * 3)(*sky), bias, bounce, shadows,
                     None, None, mode)
    if prim_rgb is not None:
        sp._rgb = _f32(prim_rgb)
        sp.prim_rgb = _p(sp._rgb)
    if prim_mat is not None:
        sp._mat = np.ascontiguousarray(prim_mat, np.uint8)
        sp.prim_mat = sp._mat.ctypes.data_as(C.c_void_p)
    return sp


def render_samples(bvh: OracleBVH, mesh, rays: dict, sp: ShadeParams, threads: int = 8) -> dict:
    if hasattr(mesh, "cx"):
        (vx, vy, vz, idx), keep = [None] * 4, None  # sphere leaves carry their centres
    else:
        (vx, vy, vz, idx), keep = mesh_ptrs(mesh)
    rs, rk = rays_struct(rays)
    n = len(rk[0])
    out = dict(rgb=np.empty(3 * n, np.float32), hit0=np.empty(n, np.int32),
               hit1=np.empty(n, np.int32), shadow0=np.empty(n, np.uint8),
               shadow1=np.empty(n, np.uint8), counts=np.zeros(8, np.uint64))
    lib().or_render_samples(bvh.h, vx, vy, vz, idx, C.byref(rs), n, C.byref(sp),
                            *(_p(out[k]) for k in ("rgb", "hit0", "hit1", "shadow0", "shadow1",
                                                   "counts")), threads)
    return out


def resolve_ppm(rgb: np.ndarray, npix: int, spp: int) -> np.ndarray:
    rgb = np.ascontiguousarray(rgb, np.float32)
    out = np.empty(3 * npix, np.uint8)
    lib().or_resolve_ppm(_p(rgb), npix, spp, _p(out))
    return out


class PortBVH:
    """CPU baseline: the reference's median-split BVH (constructBVHNew) or its
    "LBVH" (constructLBVHTree, lbvh=True) + collect-then-test query."""

    def __init__(self, mesh, lbvh: bool = False):
        (vx, vy, vz, idx), self._keep = mesh_ptrs(mesh)
        nn = C.c_uint32()
        fn = lib().port_lbvh_build if lbvh else lib().port_bvh_build
        self.h = fn(vx, vy, vz, idx, mesh.ntri, C.byref(nn))
        self.n_nodes = nn.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().port_bvh_free(self.h)
            self.h = None

    def closest(self, rays: dict, threads: int = 1) -> dict:
        rk = [_f32(rays[k]) for k in RAY_FIELDS]
        n = len(rk[0])
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        lib().port_closest(self.h, *(_p(a) for a in rk), n, _p(prim), _p(t), threads)
        return dict(prim=prim, t=t)


def ref_sphere_scene_colors(sph, rays: dict, mat=None) -> np.ndarray | None:
    """castRay of the reference itself (NONE loop, trace_more stub) over a
    sphere set, every sphere DIFFUSE_AND_GLOSSY or with materialType mat[i]:
    one colour per ray (3 floats), or None when the reference build is absent."""
    L = ref()
    if L is None:
        return None
    keep = [_f32(sph.cx), _f32(sph.cy), _f32(sph.cz), _f32(sph.r), _f32(sph.rgb)]
    rk = [_f32(rays[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
    n = len(rk[0])
    out = np.empty(3 * n, np.float32)
    m = None if mat is None else np.ascontiguousarray(mat, np.uint8)
    L.ref_sphere_scene_colors_mat(*(_p(a) for a in keep), None if m is None else _p(m),
                                  len(keep[0]), *(_p(a) for a in rk), n, _p(out))
    return out


def ref_kdtree_any(sph, rays: dict) -> np.ndarray | None:
    """The reference's KDTREE path (its SAH kd-tree and kdtreeIntersect) over a
    sphere set: 1 per ray that hits a sphere, or None without the reference build."""
    L = ref()
    if L is None:
        return None
    keep = [_f32(sph.cx), _f32(sph.cy), _f32(sph.cz), _f32(sph.r)]
    rk = [_f32(rays[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
    n = len(rk[0])
    out = np.empty(n, np.uint8)
    L.ref_kdtree_any(*(_p(a) for a in keep), len(keep[0]), *(_p(a) for a in rk), n, _p(out))
    return out
