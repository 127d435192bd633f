"""Python bindings (ctypes) for librq, the MI355X ray-query engine.

The product is the C ABI in include/rq/*.h implemented by librq.so (HIP
kernels for gfx950). This module only marshals numpy arrays through that ABI;
it has no compute path of its own, and it raises if librq.so is missing.

Reference interface mirrored (Alhajras/Raytracer-Data-structures):
  Settings.dataStructure / AccType  settings.h:7-17, accelerators.h:21
  constructLBVHTree                 accelerators.h:570-586  -> Context.build_lbvh
  boxIntersect + candidate loop     accelerators.h:668-690, main.cpp:343-358
                                                             -> Context.closest_hit
  kdtreeIntersect (any hit)         accelerators.h:997-1086 -> Context.any_hit
  render                            main.cpp:541-566         -> Renderer
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "librq.so"


class AccType(enum.IntEnum):
    """settings.h dataStructure values (accelerators.h:21) plus GPU_LBVH."""

    BVH = 0
    KDTREE = 1
    UNIFORM_GRID = 2
    LBVH = 3
    NONE = 4
    GPU_LBVH = 5


MORTON_REF_AFFINE = 0
MORTON_SCENE_AABB = 1
CAMERA_REF_JITTER = 0
CAMERA_PHILOX = 1
RENDER_SHADE = 0
RENDER_OCCLUSION = 1  # castRay's KDTREE case: black on any hit, else sky
MAT_DIFFUSE_AND_GLOSSY = 0
MAT_REFLECTION_AND_REFRACTION = 1
MAT_REFLECTION = 2

PRIM_TRIANGLE = 0
PRIM_SPHERE = 1

RQ_OK = 0
RQ_E_INVALID, RQ_E_NOMEM, RQ_E_HIP, RQ_E_IO = -1, -2, -3, -4
RQ_E_NODEVICE, RQ_E_UNSUPPORTED, RQ_E_DEVICE, RQ_E_CANCELED = -5, -6, -7, -8
ERRORS = {-1: "RQ_E_INVALID", -2: "RQ_E_NOMEM", -3: "RQ_E_HIP", -4: "RQ_E_IO",
          -5: "RQ_E_NODEVICE", -6: "RQ_E_UNSUPPORTED", -7: "RQ_E_DEVICE", -8: "RQ_E_CANCELED"}


class RQError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class NoMemory(RQError):
    pass


# ---------------------------------------------------------------- structs
class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("prim", "t", "u", "v")]


class BvhInfo(C.Structure):
    _fields_ = [("n_prims", C.c_uint32), ("n_internal", C.c_uint32), ("height", C.c_uint32),
                ("morton_mode", C.c_int32), ("root_lo", C.c_float * 3),
                ("root_hi", C.c_float * 3), ("device_bytes", C.c_size_t)]


class TraceVariant(C.Structure):
    _fields_ = [("block", C.c_int32), ("lds_depth", C.c_int32), ("tstack", C.c_int32),
                ("wide", C.c_int32), ("closest_waves", C.c_int32), ("any_waves", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("vx", C.POINTER(C.c_float)), ("vy", C.POINTER(C.c_float)),
                ("vz", C.POINTER(C.c_float)), ("nv", C.c_uint32),
                ("idx", C.POINTER(C.c_uint32)), ("ntri", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("camera", C.c_int32), ("seed", C.c_uint64), ("light", C.c_float * 3),
                ("emission", C.c_float * 3), ("surface", C.c_float * 3),
                ("sky", C.c_float * 3), ("bias", C.c_float), ("bounce", C.c_int32),
                ("shadows", C.c_int32), ("materials", C.c_int32), ("mode", C.c_int32)]


class Spheres(C.Structure):
    _fields_ = [("cx", C.POINTER(C.c_float)), ("cy", C.POINTER(C.c_float)),
                ("cz", C.POINTER(C.c_float)), ("r", C.POINTER(C.c_float)),
                ("rgb", C.POINTER(C.c_float)), ("n", C.c_uint32)]


class FrameCounts(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("bounce", C.c_uint64), ("shadow", C.c_uint64)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("nrays", C.c_uint32), ("rays", Rays), ("hits", Hits),
                ("occluded", C.c_void_p)]


class BatchConfig(C.Structure):
    _fields_ = [("lds_depth", C.c_int32), ("tstack", C.c_int32), ("ray_tile", C.c_uint32),
                ("refill_idle", C.c_uint32)]


class Pinhole(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("width", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("new_mesh", C.POINTER(Mesh)), ("nrays", C.c_uint32),
                ("source", C.c_int32), ("host_rays", Rays), ("pinhole", Pinhole),
                ("lo", C.c_float * 3), ("hi", C.c_float * 3), ("seed", C.c_uint64),
                ("query", C.c_int32), ("prim", C.c_void_p), ("t", C.c_void_p),
                ("occluded", C.c_void_p)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("mesh", C.c_uint32), ("steps", C.c_uint32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("mesh_slots", C.c_uint32), ("slot_tris", C.c_uint32),
                ("arena_bytes", C.c_size_t), ("max_inflight", C.c_uint32),
                ("step_rays", C.c_uint32), ("chunk_rays", C.c_uint32), ("ray_tile", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("steps", C.c_uint64), ("rays", C.c_uint64), ("builds", C.c_uint64),
                ("max_step_requests", C.c_uint32), ("resident", C.c_uint32),
                ("arena_peak", C.c_size_t), ("busy_s", C.c_double)]


RAYS_HOST, RAYS_PINHOLE, RAYS_UNIFORM = 0, 1, 2
Q_CLOSEST, Q_ANY = 1, 2
REQ_QUEUED, REQ_ADMITTED, REQ_DONE = 0, 1, 2

BATCH_PHASES = ("init", "bounds", "morton", "sort0", "sort1", "sort2", "sort3", "karras", "refit",
                "refit_top", "widen", "trace")


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run raytracer-data-structures_amd/build.py")
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, u64, i32, f32p = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.POINTER(C.c_float)
    sig = {
        "rq_create": (i32, [C.POINTER(vp), i32, C.c_size_t]),
        "rq_destroy": (i32, [vp]),
        "rq_last_error": (C.c_char_p, [vp]),
        "rq_version": (C.c_char_p, []),
        "rq_bytes_in_use": (C.c_size_t, [vp]),
        "rq_scene_upload": (i32, [vp, vp, vp, vp, u32, vp, u32, i32, C.POINTER(vp), vp]),
        "rq_scene_update_vertices": (i32, [vp, vp, vp, vp, vp, vp]),
        "rq_build_lbvh": (i32, [vp, vp, i32, C.POINTER(vp), vp]),
        "rq_rebuild_lbvh": (i32, [vp, vp, vp]),
        "rq_bvh_treelets": (i32, [vp, vp, i32, vp]),
        "rq_bvh_info_get": (i32, [vp, vp, C.POINTER(BvhInfo)]),
        "rq_bvh_export": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "rq_bvh_export_wide": (i32, [vp, vp, vp, vp]),
        "rq_closest_hit": (i32, [vp, vp, C.POINTER(Rays), u32, C.POINTER(Hits), vp]),
        "rq_any_hit": (i32, [vp, vp, C.POINTER(Rays), u32, vp, vp]),
        "rq_closest_hit_host": (i32, [vp, vp, C.POINTER(Rays), u32, C.POINTER(Hits), vp]),
        "rq_any_hit_host": (i32, [vp, vp, C.POINTER(Rays), u32, vp, vp]),
        "rq_cast_ray": (i32, [vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                              C.c_float, C.POINTER(C.c_int32), C.POINTER(C.c_float),
                              C.POINTER(C.c_float)]),
        "rq_occluded_ray": (i32, [vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                  C.c_float, C.POINTER(C.c_uint8)]),
        "rq_closest_hit_counted": (i32, [vp, vp, C.POINTER(Rays), u32, C.POINTER(Hits), vp, vp,
                                         vp]),
        "rq_any_hit_counted": (i32, [vp, vp, C.POINTER(Rays), u32, vp, vp, vp, vp]),
        "rq_trace_variant_count": (i32, []),
        "rq_trace_variant_get": (i32, [i32, C.POINTER(TraceVariant)]),
        "rq_trace_config_set": (i32, [vp, i32, i32, u32]),
        "rq_trace_config_get": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(u32)]),
        "rq_trace_refill_set": (i32, [vp, u32, u32]),
        "rq_trace_refill_get": (i32, [vp, C.POINTER(u32), C.POINTER(u32)]),
        "rq_trace_slice_set": (i32, [vp, u32, u32]),
        "rq_trace_slice_get": (i32, [vp, C.POINTER(u32), C.POINTER(u32)]),
        "rq_autotune_trace": (i32, [vp, vp, C.POINTER(Rays), u32, i32, i32, C.POINTER(i32),
                                    C.POINTER(u32), C.POINTER(C.c_double)]),
        "rq_device_alloc": (i32, [vp, C.c_size_t, C.POINTER(vp)]),
        "rq_device_free": (i32, [vp, vp]),
        "rq_memcpy_h2d": (i32, [vp, vp, vp, C.c_size_t, vp]),
        "rq_memcpy_d2h": (i32, [vp, vp, vp, C.c_size_t, vp]),
        "rq_memcpy_d2d": (i32, [vp, vp, vp, C.c_size_t, vp]),
        "rq_memset_d": (i32, [vp, vp, i32, C.c_size_t, vp]),
        "rq_sync": (i32, [vp, vp]),
        "rq_device_errors": (i32, [vp, C.POINTER(u32)]),
        "rq_stream_create": (i32, [vp, C.POINTER(vp)]),
        "rq_stream_destroy": (i32, [vp, vp]),
        "rq_stream_order": (i32, [vp, vp, vp]),
        "rq_stream_forget": (i32, [vp, vp]),
        "rq_profile_enable": (i32, [vp, i32]),
        "rq_profile_reset": (i32, [vp]),
        "rq_profile_only": (i32, [vp, C.c_char_p]),
        "rq_profile_read": (i32, [vp, vp, vp, vp, i32]),
        "rq_release": (i32, [vp, vp]),
        "rq_gen_icosphere": (i32, [u32, u32, C.c_float, C.c_float, f32p, C.c_float,
                                   C.POINTER(Mesh)]),
        "rq_gen_icosphere_field": (i32, [u32, u32, u32, u64, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.POINTER(Mesh)]),
        "rq_gen_soup": (i32, [u32, u64, C.c_float, C.c_float, C.c_float, C.POINTER(Mesh)]),
        "rq_obj_read": (i32, [C.c_char_p, C.POINTER(Mesh)]),
        "rq_obj_read_ex": (i32, [C.c_char_p, i32, C.POINTER(Mesh)]),
        "rq_obj_write": (i32, [C.c_char_p, C.POINTER(Mesh)]),
        "rq_mesh_bounds": (None, [C.POINTER(Mesh), f32p, f32p]),
        "rq_mesh_free": (None, [C.POINTER(Mesh)]),
        "rq_render_defaults": (None, [C.POINTER(RenderParams)]),
        "rq_renderer_create": (i32, [vp, C.POINTER(RenderParams), C.POINTER(vp)]),
        "rq_renderer_destroy": (i32, [vp, vp]),
        "rq_render_frame": (i32, [vp, vp, vp, vp]),
        "rq_render_wait_traced": (i32, [vp, vp, vp]),
        "rq_render_counts": (i32, [vp, vp, C.POINTER(FrameCounts)]),
        "rq_render_read_image": (i32, [vp, vp, vp, vp]),
        "rq_render_read_samples": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "rq_render_primary_rays": (i32, [vp, vp, C.POINTER(Rays), C.POINTER(u32)]),
        "rq_render_rays": (i32, [vp, vp, i32, C.POINTER(Rays), C.POINTER(u32)]),
        "rq_raygen_camera": (i32, [vp, C.POINTER(RenderParams), vp, vp, vp, vp, vp, vp, vp, vp,
                                   vp, vp]),
        "rq_raygen_uniform": (i32, [vp, u32, u64, f32p, f32p, vp, vp, vp, vp, vp, vp, vp, vp,
                                    vp]),
        "rq_write_ppm": (i32, [C.c_char_p, u32, u32, vp]),
        "rq_scene_upload_spheres": (i32, [vp, vp, vp, vp, vp, vp, u32, C.POINTER(vp), vp]),
        "rq_scene_set_materials": (i32, [vp, vp, vp, vp]),
        "rq_bvh_blob_size": (i32, [vp, vp, C.POINTER(C.c_size_t)]),
        "rq_bvh_save": (i32, [vp, vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "rq_bvh_load": (i32, [vp, vp, vp, C.c_size_t, C.POINTER(vp), vp]),
        "rq_bvh_save_file": (i32, [vp, vp, C.c_char_p]),
        "rq_bvh_load_file": (i32, [vp, vp, C.c_char_p, C.POINTER(vp)]),
        "rq_bvh_blob_validate": (i32, [vp, C.c_size_t, i32, vp, vp, vp, u32, vp, vp, u32,
                                       C.c_char_p, C.c_size_t]),
        "rq_gen_reference_scene": (i32, [C.c_char_p, u32, u32, C.POINTER(Spheres)]),
        "rq_spheres_free": (None, [C.POINTER(Spheres)]),
        "rq_batch_create": (i32, [vp, C.POINTER(BatchItem), u32, i32, C.POINTER(vp)]),
        "rq_batch_destroy": (i32, [vp, vp]),
        "rq_batch_run": (i32, [vp, vp, vp]),
        "rq_batch_run_graph": (i32, [vp, vp, vp]),
        "rq_batch_bvh": (i32, [vp, vp, u32, C.POINTER(vp)]),
        "rq_batch_config_count": (i32, []),
        "rq_batch_config_get": (i32, [i32, C.POINTER(BatchConfig)]),
        "rq_batch_set_config": (i32, [vp, vp, i32, u32, u32]),
        "rq_batch_autotune": (i32, [vp, vp, i32, C.POINTER(i32), C.POINTER(u32), C.POINTER(u32),
                                    C.POINTER(C.c_double)]),
        "rq_batch_phase_items": (i32, [vp, vp, vp, i32]),
        "rq_batch_phase_ms": (i32, [vp, vp, vp, i32]),
        "rq_batch_errors": (i32, [vp, vp, C.POINTER(u32)]),
        "rq_batch_occupancy": (i32, [vp, vp, C.POINTER(u32), C.POINTER(u32)]),
        "rq_batch_set_build_waves": (i32, [vp, vp, i32]),
        "rq_batch_get_config": (i32, [vp, vp, C.POINTER(C.c_int), C.POINTER(u32), C.POINTER(u32)]),
        "rq_batch_get_build_waves": (i32, [vp, vp, C.POINTER(C.c_int)]),
        "rq_server_defaults": (None, [C.POINTER(ServerConfig)]),
        "rq_server_create": (i32, [i32, C.POINTER(ServerConfig), C.POINTER(vp)]),
        "rq_server_destroy": (i32, [vp]),
        "rq_server_load_mesh": (i32, [vp, C.POINTER(Mesh), C.POINTER(u32)]),
        "rq_server_submit": (i32, [vp, C.POINTER(Request), C.POINTER(u64)]),
        "rq_server_poll": (i32, [vp, u64, C.POINTER(RequestStatus)]),
        "rq_server_wait": (i32, [vp, u64, C.POINTER(RequestStatus)]),
        "rq_server_forget": (i32, [vp, u64]),
        "rq_server_cancel": (i32, [vp, u64]),
        "rq_server_take_done": (i32, [vp, vp, vp, u32, C.POINTER(u32)]),
        "rq_server_stats_get": (i32, [vp, C.POINTER(ServerStats)]),
        "rq_server_last_error": (C.c_char_p, [vp]),
        "rq_server_clock": (C.c_double, []),
        "rq_host_alloc": (vp, [C.c_size_t]),
        "rq_host_free": (None, [vp]),
        "rq_raygen_pinhole": (i32, [vp, C.POINTER(Pinhole), u32, u64, vp, vp, vp, vp, vp, vp, vp,
                                    vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- meshes (host)
@dataclass
class MeshArrays:
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    idx: np.ndarray  # (ntri*3,) uint32

    @property
    def nv(self) -> int:
        return int(self.vx.size)

    @property
    def ntri(self) -> int:
        return int(self.idx.size // 3)

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        v = np.stack([self.vx, self.vy, self.vz])
        return v.min(axis=1), v.max(axis=1)


@dataclass
class SphereArrays:
    """Spheres: centres, radii, surface colour (3 per sphere)."""
    cx: np.ndarray
    cy: np.ndarray
    cz: np.ndarray
    r: np.ndarray
    rgb: np.ndarray

    @property
    def n(self) -> int:
        return int(self.cx.size)

    @property
    def ntri(self) -> int:  # primitives, as for meshes
        return self.n

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        c = np.stack([self.cx, self.cy, self.cz])
        return (c - self.r).min(axis=1), (c + self.r).max(axis=1)


def reference_scene(obj_path: str | Path, clones: int = 1, max_vertices: int = 0) -> SphereArrays:
    """The reference's sphere scene from an OBJ file (rq_gen_reference_scene)."""
    s = Spheres()
    _check_gen(lib().rq_gen_reference_scene(str(obj_path).encode(), clones, max_vertices,
                                            C.byref(s)), f"reference_scene {obj_path}")
    try:
        n = s.n
        arr = lambda p, k: np.ctypeslib.as_array(p, (k,)).copy()  # noqa: E731
        out = SphereArrays(arr(s.cx, n), arr(s.cy, n), arr(s.cz, n), arr(s.r, n),
                           arr(s.rgb, 3 * n))
    finally:
        lib().rq_spheres_free(C.byref(s))
    return out


def _take_mesh(m: Mesh) -> MeshArrays:
    L = lib()
    try:
        out = MeshArrays(
            np.ctypeslib.as_array(m.vx, (m.nv,)).copy() if m.nv else np.zeros(0, np.float32),
            np.ctypeslib.as_array(m.vy, (m.nv,)).copy() if m.nv else np.zeros(0, np.float32),
            np.ctypeslib.as_array(m.vz, (m.nv,)).copy() if m.nv else np.zeros(0, np.float32),
            np.ctypeslib.as_array(m.idx, (3 * m.ntri,)).copy() if m.ntri
            else np.zeros(0, np.uint32))
    finally:
        L.rq_mesh_free(C.byref(m))
    return out


def _check_gen(rc: int, what: str) -> None:
    if rc != RQ_OK:
        raise RQError(rc, what)


def icosphere(level: int, seed: int, amp: float = 0.1, freq: float = 4.0,
              center=(0.0, 0.0, 0.0), scale: float = 1.0) -> MeshArrays:
    m = Mesh()
    c = (C.c_float * 3)(*center)
    _check_gen(lib().rq_gen_icosphere(level, seed, amp, freq, c, scale, C.byref(m)), "icosphere")
    return _take_mesh(m)


def icosphere_field(count: int, level: int, seed0: int, placement_seed: int, box: float,
                    rmin: float, amp: float = 0.1, freq: float = 4.0) -> MeshArrays:
    m = Mesh()
    _check_gen(lib().rq_gen_icosphere_field(count, level, seed0, placement_seed, box, rmin, amp,
                                            freq, C.byref(m)), "icosphere_field")
    return _take_mesh(m)


def soup(ntri: int, seed: int, mean_edge=0.02, sd_edge=0.005, min_edge=0.002) -> MeshArrays:
    m = Mesh()
    _check_gen(lib().rq_gen_soup(ntri, seed, mean_edge, sd_edge, min_edge, C.byref(m)), "soup")
    return _take_mesh(m)


def obj_read(path: str | Path, threads: int = 0) -> MeshArrays:
    """Parallel memory-mapped OBJ reader (threads 0 = automatic, 1 = sequential)."""
    m = Mesh()
    _check_gen(lib().rq_obj_read_ex(str(path).encode(), threads, C.byref(m)), f"obj_read {path}")
    return _take_mesh(m)


def obj_write(path: str | Path, mesh: MeshArrays) -> None:
    vx, vy, vz = (np.ascontiguousarray(a, np.float32) for a in (mesh.vx, mesh.vy, mesh.vz))
    idx = np.ascontiguousarray(mesh.idx, np.uint32)
    m = Mesh(vx.ctypes.data_as(C.POINTER(C.c_float)), vy.ctypes.data_as(C.POINTER(C.c_float)),
             vz.ctypes.data_as(C.POINTER(C.c_float)), vx.size,
             idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size // 3)
    _check_gen(lib().rq_obj_write(str(path).encode(), C.byref(m)), f"obj_write {path}")


def render_defaults() -> RenderParams:
    p = RenderParams()
    lib().rq_render_defaults(C.byref(p))
    return p


# ---------------------------------------------------------------- device
RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


class DeviceArray:
    """A device allocation owned by a Context (charged to its HBM budget)."""

    owned = True

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx._check(lib().rq_device_alloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            self.ctx._check(lib().rq_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), a.nbytes, stream))
        return self

    def download(self, dtype, count: int, stream=None) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            self.ctx._check(lib().rq_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes,
                                                stream))
        return out

    def fill(self, byte: int, stream=None) -> "DeviceArray":
        self.ctx._check(lib().rq_memset_d(self.ctx.h, self.ptr, byte, self.nbytes, stream))
        return self

    def free(self) -> None:
        if self.ptr and self.owned:
            lib().rq_device_free(self.ctx.h, self.ptr)
        self.ptr = None


class DeviceRays:
    """Device SoA ray batch (8 float32 arrays)."""

    def __init__(self, ctx: "Context", n: int):
        self.n = int(n)
        self.arrays = {f: DeviceArray(ctx, 4 * self.n) for f in RAY_FIELDS}

    @classmethod
    def from_host(cls, ctx: "Context", rays: dict) -> "DeviceRays":
        n = len(rays["ox"])
        r = cls(ctx, n)
        for f in RAY_FIELDS:
            r.arrays[f].upload(np.asarray(rays[f], np.float32))
        return r

    def struct(self) -> Rays:
        return Rays(*(self.arrays[f].ptr for f in RAY_FIELDS))

    def download(self) -> dict:
        return {f: self.arrays[f].download(np.float32, self.n) for f in RAY_FIELDS}

    def free(self) -> None:
        for a in self.arrays.values():
            a.free()


class Context:
    """rq_ctx: one device, an HBM budget, a host thread."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        rc = lib().rq_create(C.byref(h), device, hbm_budget)
        if rc != RQ_OK:
            raise RQError(rc, f"rq_create(device={device}) failed (no gfx950 device?)")
        self.h = h.value

    def close(self) -> None:
        if self.h:
            lib().rq_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != RQ_OK:
            msg = lib().rq_last_error(self.h).decode()
            raise (NoMemory if rc == -2 else RQError)(rc, msg)

    @property
    def bytes_in_use(self) -> int:
        return int(lib().rq_bytes_in_use(self.h))

    # scene + build
    def upload_scene(self, mesh: MeshArrays, stream=None) -> int:
        vx, vy, vz = (np.ascontiguousarray(a, np.float32) for a in (mesh.vx, mesh.vy, mesh.vz))
        idx = np.ascontiguousarray(mesh.idx, np.uint32)
        s = C.c_void_p()
        self._check(lib().rq_scene_upload(self.h, _ptr(vx), _ptr(vy), _ptr(vz), vx.size,
                                          _ptr(idx), idx.size // 3, 0, C.byref(s), stream))
        return s.value

    def upload_spheres(self, sph: SphereArrays, colors: bool = True, stream=None) -> int:
        a = [np.ascontiguousarray(x, np.float32) for x in (sph.cx, sph.cy, sph.cz, sph.r, sph.rgb)]
        s = C.c_void_p()
        self._check(lib().rq_scene_upload_spheres(self.h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
                                                  _ptr(a[3]), _ptr(a[4]) if colors else None,
                                                  a[0].size, C.byref(s), stream))
        return s.value

    def set_materials(self, scene: int, mat, stream=None) -> None:
        """Material per primitive (MAT_*), read by renderers with materials = 1."""
        m = np.ascontiguousarray(mat, np.uint8)
        self._check(lib().rq_scene_set_materials(self.h, scene, _ptr(m), stream))

    def build_lbvh(self, scene: int, morton_mode: int = MORTON_SCENE_AABB, stream=None) -> int:
        b = C.c_void_p()
        self._check(lib().rq_build_lbvh(self.h, scene, morton_mode, C.byref(b), stream))
        return b.value

    def rebuild(self, bvh: int, stream=None) -> None:
        self._check(lib().rq_rebuild_lbvh(self.h, bvh, stream))

    def treelets(self, bvh: int, rounds: int = 1, stream=None) -> None:
        """Treelet restructuring of a built BVH in place (rq_bvh_treelets)."""
        self._check(lib().rq_bvh_treelets(self.h, bvh, rounds, stream))

    def bvh_info(self, bvh: int) -> BvhInfo:
        info = BvhInfo()
        self._check(lib().rq_bvh_info_get(self.h, bvh, C.byref(info)))
        return info

    def export(self, bvh: int, n: int) -> dict:
        codes = np.empty(n, np.uint32)
        sidx = np.empty(n, np.uint32)
        left = np.empty(max(n - 1, 0), np.int32)
        right = np.empty(max(n - 1, 0), np.int32)
        boxes = np.empty(max(2 * n - 1, 0) * 6, np.float32)
        self._check(lib().rq_bvh_export(self.h, bvh, _ptr(codes), _ptr(sidx), _ptr(left),
                                        _ptr(right), _ptr(boxes)))
        return dict(codes=codes, sorted_idx=sidx, left=left, right=right,
                    boxes=boxes.reshape(-1, 6))

    def export_wide(self, bvh: int, n: int) -> dict:
        """The 4-wide nodes (rq_bvh_export_wide): boxes[n-1, 4, 6] and
        links[n-1, 4], defined for the root and nodes at even depth."""
        m = max(n - 1, 0)
        boxes = np.zeros(m * 24, np.float32)
        links = np.zeros(m * 4, np.int32)
        self._check(lib().rq_bvh_export_wide(self.h, bvh, _ptr(boxes), _ptr(links)))
        return dict(boxes=boxes.reshape(m, 4, 6), links=links.reshape(m, 4))

    def bvh_save(self, bvh: int) -> bytes:
        """The built LBVH as a blob (rq_bvh_save)."""
        n = C.c_size_t()
        self._check(lib().rq_bvh_blob_size(self.h, bvh, C.byref(n)))
        buf = (C.c_char * n.value)()
        self._check(lib().rq_bvh_save(self.h, bvh, buf, n.value, None))
        return bytes(buf)

    def bvh_load(self, scene: int, blob: bytes) -> int:
        b = C.c_void_p()
        self._check(lib().rq_bvh_load(self.h, scene, blob, len(blob), C.byref(b), None))
        return b.value

    def bvh_save_file(self, bvh: int, path) -> None:
        self._check(lib().rq_bvh_save_file(self.h, bvh, str(path).encode()))

    def bvh_load_file(self, scene: int, path) -> int:
        b = C.c_void_p()
        self._check(lib().rq_bvh_load_file(self.h, scene, str(path).encode(), C.byref(b)))
        return b.value

    def release(self, handle: int) -> None:
        self._check(lib().rq_release(self.h, handle))

    # queries (device buffers)
    def closest_hit_dev(self, bvh: int, rays: DeviceRays, hits: dict, counted=None,
                        stream=None) -> None:
        r = rays.struct()
        h = Hits(hits["prim"].ptr, hits["t"].ptr if "t" in hits else None,
                 hits["u"].ptr if "u" in hits else None, hits["v"].ptr if "v" in hits else None)
        if counted:
            self._check(lib().rq_closest_hit_counted(self.h, bvh, C.byref(r), rays.n, C.byref(h),
                                                     counted[0].ptr, counted[1].ptr, stream))
        else:
            self._check(lib().rq_closest_hit(self.h, bvh, C.byref(r), rays.n, C.byref(h), stream))

    def any_hit_dev(self, bvh: int, rays: DeviceRays, occ: DeviceArray, counted=None,
                    stream=None) -> None:
        r = rays.struct()
        if counted:
            self._check(lib().rq_any_hit_counted(self.h, bvh, C.byref(r), rays.n, occ.ptr,
                                                 counted[0].ptr, counted[1].ptr, stream))
        else:
            self._check(lib().rq_any_hit(self.h, bvh, C.byref(r), rays.n, occ.ptr, stream))

    # queries (host round trip, for tests)
    def closest_hit(self, bvh: int, rays: dict, counted: bool = False) -> dict:
        n = len(rays["ox"])
        dr = DeviceRays.from_host(self, rays)
        hits = {k: DeviceArray(self, 4 * max(n, 1)) for k in ("prim", "t", "u", "v")}
        cnt = (DeviceArray(self, 4 * max(n, 1)), DeviceArray(self, 4 * max(n, 1))) \
            if counted else None
        try:
            self.closest_hit_dev(None if bvh is None else bvh, dr, hits, cnt)
            self.sync()
            out = {"prim": hits["prim"].download(np.int32, n)}
            for k in ("t", "u", "v"):
                out[k] = hits[k].download(np.float32, n)
            if counted:
                out["visits"] = cnt[0].download(np.uint32, n)
                out["tests"] = cnt[1].download(np.uint32, n)
            return out
        finally:
            dr.free()
            for a in hits.values():
                a.free()
            if cnt:
                cnt[0].free(), cnt[1].free()

    def any_hit(self, bvh: int, rays: dict, counted: bool = False) -> dict:
        n = len(rays["ox"])
        dr = DeviceRays.from_host(self, rays)
        occ = DeviceArray(self, max(n, 1))
        cnt = (DeviceArray(self, 4 * max(n, 1)), DeviceArray(self, 4 * max(n, 1))) \
            if counted else None
        try:
            self.any_hit_dev(bvh, dr, occ, cnt)
            self.sync()
            out = {"occ": occ.download(np.uint8, n)}
            if counted:
                out["visits"] = cnt[0].download(np.uint32, n)
                out["tests"] = cnt[1].download(np.uint32, n)
            return out
        finally:
            dr.free()
            occ.free()
            if cnt:
                cnt[0].free(), cnt[1].free()

    def closest_hit_host(self, bvh: int, rays: dict, uv: bool = True) -> dict:
        """rq_closest_hit_host: rays and results in host memory (staged by the library)."""
        n = len(rays["ox"])
        cols = {k: np.ascontiguousarray(rays[k], np.float32) for k in RAY_FIELDS}
        hr = Rays(*[_ptr(cols[k]) for k in RAY_FIELDS])
        out = {"prim": np.empty(n, np.int32), "t": np.empty(n, np.float32)}
        if uv:
            out["u"], out["v"] = np.empty(n, np.float32), np.empty(n, np.float32)
        h = Hits(_ptr(out["prim"]), _ptr(out["t"]), _ptr(out.get("u")), _ptr(out.get("v")))
        self._check(lib().rq_closest_hit_host(self.h, bvh, C.byref(hr), n, C.byref(h), None))
        return out

    def any_hit_host(self, bvh: int, rays: dict) -> np.ndarray:
        """rq_any_hit_host: occlusion flags of host rays."""
        n = len(rays["ox"])
        cols = {k: np.ascontiguousarray(rays[k], np.float32) for k in RAY_FIELDS}
        hr = Rays(*[_ptr(cols[k]) for k in RAY_FIELDS])
        occ = np.empty(n, np.uint8)
        self._check(lib().rq_any_hit_host(self.h, bvh, C.byref(hr), n, _ptr(occ), None))
        return occ

    def cast_ray(self, bvh: int, orig, d, tmin: float = 0.0, tmax: float = float("inf")):
        """rq_cast_ray: (prim, t, (u, v)) of one host ray."""
        o3, d3 = (C.c_float * 3)(*orig), (C.c_float * 3)(*d)
        prim, t, uv = C.c_int32(), C.c_float(), (C.c_float * 2)()
        self._check(lib().rq_cast_ray(self.h, bvh, o3, d3, tmin, tmax, C.byref(prim),
                                      C.byref(t), uv))
        return prim.value, t.value, (uv[0], uv[1])

    def occluded_ray(self, bvh: int, orig, d, tmin: float = 0.0,
                     tmax: float = float("inf")) -> bool:
        """rq_occluded_ray: whether one host ray hits anything in (tmin, tmax)."""
        o3, d3 = (C.c_float * 3)(*orig), (C.c_float * 3)(*d)
        occ = C.c_uint8()
        self._check(lib().rq_occluded_ray(self.h, bvh, o3, d3, tmin, tmax, C.byref(occ)))
        return bool(occ.value)

    def raygen_uniform(self, n: int, seed: int, lo, hi, stream=None) -> DeviceRays:
        r = DeviceRays(self, n)
        a = r.arrays
        lo3 = (C.c_float * 3)(*lo)
        hi3 = (C.c_float * 3)(*hi)
        self._check(lib().rq_raygen_uniform(self.h, n, seed, lo3, hi3,
                                            *(a[f].ptr for f in RAY_FIELDS), stream))
        return r

    def raygen_pinhole(self, n: int, seed: int, eye, look_at, up=(0.0, 1.0, 0.0),
                       fov_deg: float = 40.0, width: int = 1024, stream=None) -> DeviceRays:
        r = DeviceRays(self, n)
        a = r.arrays
        p = Pinhole((C.c_float * 3)(*eye), (C.c_float * 3)(*look_at), (C.c_float * 3)(*up),
                    fov_deg, width)
        self._check(lib().rq_raygen_pinhole(self.h, C.byref(p), n, seed,
                                            *(a[f].ptr for f in RAY_FIELDS), stream))
        return r

    def raygen_camera(self, params: RenderParams, jitter: np.ndarray | None = None,
                      stream=None) -> DeviceRays:
        n = params.width * params.height * params.spp
        r = DeviceRays(self, n)
        a = r.arrays
        jd = None
        if jitter is not None:
            jd = DeviceArray(self, 8 * jitter.size).upload(np.asarray(jitter, np.float64))
        try:
            self._check(lib().rq_raygen_camera(self.h, C.byref(params), jd.ptr if jd else None,
                                               *(a[f].ptr for f in RAY_FIELDS), stream))
            self.sync()
        finally:
            if jd:
                jd.free()
        return r

    def copy(self, dst: "DeviceArray", src: "DeviceArray", nbytes: int, stream=None) -> None:
        self._check(lib().rq_memcpy_d2d(self.h, dst.ptr, src.ptr, nbytes, stream))

    def sync(self, stream=None) -> None:
        self._check(lib().rq_sync(self.h, stream))

    def device_errors(self) -> int:
        bits = C.c_uint32()
        self._check(lib().rq_device_errors(self.h, C.byref(bits)))
        return bits.value

    def stream_create(self):
        s = C.c_void_p()
        self._check(lib().rq_stream_create(self.h, C.byref(s)))
        return s.value

    def stream_destroy(self, s) -> None:
        self._check(lib().rq_stream_destroy(self.h, s))

    def stream_order(self, waiter, signaler) -> None:
        """rq_stream_order: waiter runs after the work enqueued on signaler."""
        self._check(lib().rq_stream_order(self.h, waiter, signaler))

    def stream_forget(self, s) -> None:
        """rq_stream_forget: free the overflow area of a caller-owned stream."""
        self._check(lib().rq_stream_forget(self.h, s))

    # traversal variants + autotuner
    @staticmethod
    def trace_variants() -> list[dict]:
        out = []
        for i in range(lib().rq_trace_variant_count()):
            v = TraceVariant()
            lib().rq_trace_variant_get(i, C.byref(v))
            out.append(dict(index=i, block=v.block, lds_depth=v.lds_depth, tstack=v.tstack,
                            wide=v.wide, closest_waves=v.closest_waves, any_waves=v.any_waves))
        return out

    def trace_config(self, closest: int | None = None, any_hit: int | None = None,
                     refill_idle: int | None = None) -> tuple[int, int, int]:
        c, a, r = C.c_int(), C.c_int(), C.c_uint32()
        self._check(lib().rq_trace_config_get(self.h, C.byref(c), C.byref(a), C.byref(r)))
        nc = c.value if closest is None else closest
        na = a.value if any_hit is None else any_hit
        rc, ra = self.trace_refill()
        nr = r.value if refill_idle is None else refill_idle
        self._check(lib().rq_trace_config_set(self.h, nc, na, nr))
        if refill_idle is None:
            self.trace_refill(rc, ra)  # keep the per-kind thresholds
        return nc, na, nr

    def trace_refill(self, closest: int | None = None,
                     any_hit: int | None = None) -> tuple[int, int]:
        """Refill thresholds per query kind (rq_trace_refill_set / _get)."""
        c, a = C.c_uint32(), C.c_uint32()
        self._check(lib().rq_trace_refill_get(self.h, C.byref(c), C.byref(a)))
        nc = c.value if closest is None else closest
        na = a.value if any_hit is None else any_hit
        self._check(lib().rq_trace_refill_set(self.h, nc, na))
        return nc, na

    def trace_slice(self, closest: int | None = None,
                    any_hit: int | None = None) -> tuple[int, int]:
        """Largest ray-stream slices per query kind (rq_trace_slice_set / _get)."""
        c, a = C.c_uint32(), C.c_uint32()
        self._check(lib().rq_trace_slice_get(self.h, C.byref(c), C.byref(a)))
        nc = c.value if closest is None else closest
        na = a.value if any_hit is None else any_hit
        self._check(lib().rq_trace_slice_set(self.h, nc, na))
        return nc, na

    def autotune(self, bvh: int, rays: DeviceRays, any_hit: bool = False,
                 repeats: int = 3) -> tuple[int, int, float]:
        v, rf, ms = C.c_int(), C.c_uint32(), C.c_double()
        r = rays.struct()
        self._check(lib().rq_autotune_trace(self.h, bvh, C.byref(r), rays.n, 1 if any_hit else 0,
                                            repeats, C.byref(v), C.byref(rf), C.byref(ms)))
        return v.value, rf.value, ms.value

    # per-kernel timing
    def profile(self, on: bool = True) -> None:
        self._check(lib().rq_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self) -> None:
        self._check(lib().rq_profile_reset(self.h))

    def profile_only(self, names=None) -> None:
        """Time only these kernel names (None: every kernel)."""
        arg = ",".join(names).encode() if names else None
        self._check(lib().rq_profile_only(self.h, arg))

    def profile_read(self) -> dict:
        cap = 64
        names = (C.c_char * 48 * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_uint64 * cap)()
        k = lib().rq_profile_read(self.h, names, ms, cnt, cap)
        return {names[i].value.decode(): (ms[i], cnt[i]) for i in range(min(k, cap))}


class Batch:
    """rq_batch: many meshes built and traced in one persistent launch.

    items: dicts with keys scene (rq_scene handle), rays (DeviceRays or None),
    hits (dict of DeviceArray: prim required, t/u/v optional; None skips the
    closest-hit query) and occ (DeviceArray of uint8 or None). The buffers are
    kept referenced here and must not be freed before close()."""

    def __init__(self, ctx: Context, items: list[dict], morton_mode: int = MORTON_SCENE_AABB):
        self.ctx = ctx
        self.items = list(items)
        arr = (BatchItem * max(len(self.items), 1))()
        for i, it in enumerate(self.items):
            rays = it.get("rays")
            hits = it.get("hits")
            occ = it.get("occ")
            arr[i].scene = it["scene"]
            arr[i].nrays = rays.n if rays is not None else 0
            if rays is not None:
                arr[i].rays = rays.struct()
            if hits is not None:
                arr[i].hits = Hits(hits["prim"].ptr, hits["t"].ptr if "t" in hits else None,
                                   hits["u"].ptr if "u" in hits else None,
                                   hits["v"].ptr if "v" in hits else None)
            arr[i].occluded = occ.ptr if occ is not None else None
        h = C.c_void_p()
        ctx._check(lib().rq_batch_create(ctx.h, arr, len(self.items), morton_mode, C.byref(h)))
        self.h = h.value

    def close(self) -> None:
        if self.h:
            self.ctx._check(lib().rq_batch_destroy(self.ctx.h, self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, stream=None) -> None:
        self.ctx._check(lib().rq_batch_run(self.ctx.h, self.h, stream))

    def run_graph(self, stream=None) -> None:
        self.ctx._check(lib().rq_batch_run_graph(self.ctx.h, self.h, stream))

    def bvh(self, i: int) -> int:
        b = C.c_void_p()
        self.ctx._check(lib().rq_batch_bvh(self.ctx.h, self.h, i, C.byref(b)))
        return b.value

    @staticmethod
    def configs() -> list[dict]:
        out = []
        for i in range(lib().rq_batch_config_count()):
            c = BatchConfig()
            lib().rq_batch_config_get(i, C.byref(c))
            out.append(dict(index=i, lds_depth=c.lds_depth, tstack=c.tstack))
        return out

    def set_config(self, index: int, ray_tile: int, refill_idle: int) -> None:
        self.ctx._check(lib().rq_batch_set_config(self.ctx.h, self.h, index, ray_tile,
                                                  refill_idle))

    def autotune(self, repeats: int = 2) -> tuple[int, int, int, float]:
        v, t, r, ms = C.c_int(), C.c_uint32(), C.c_uint32(), C.c_double()
        self.ctx._check(lib().rq_batch_autotune(self.ctx.h, self.h, repeats, C.byref(v),
                                                C.byref(t), C.byref(r), C.byref(ms)))
        return v.value, t.value, r.value, ms.value

    def config(self) -> tuple[int, int, int]:
        """rq_batch_get_config: (trace configuration index, ray tile, refill threshold)."""
        i, t, r = C.c_int(), C.c_uint32(), C.c_uint32()
        self.ctx._check(lib().rq_batch_get_config(self.ctx.h, self.h, C.byref(i), C.byref(t),
                                                  C.byref(r)))
        return i.value, t.value, r.value

    def build_waves(self, waves: int | None = None) -> int:
        """rq_batch_set/get_build_waves: the build kernel's waves per SIMD."""
        if waves is not None:
            self.ctx._check(lib().rq_batch_set_build_waves(self.ctx.h, self.h, waves))
        w = C.c_int()
        self.ctx._check(lib().rq_batch_get_build_waves(self.ctx.h, self.h, C.byref(w)))
        return w.value

    def occupancy(self) -> tuple[int, int]:
        """rq_batch_occupancy: (blocks that took part in the last run, launch grid)."""
        m, g = C.c_uint32(), C.c_uint32()
        self.ctx._check(lib().rq_batch_occupancy(self.ctx.h, self.h, C.byref(m), C.byref(g)))
        return m.value, g.value

    def phase_items(self) -> dict:
        counts = np.zeros(len(BATCH_PHASES), np.uint32)
        lib().rq_batch_phase_items(self.ctx.h, self.h, _ptr(counts), counts.size)
        return dict(zip(BATCH_PHASES, counts.tolist()))

    def phase_ms(self) -> dict:
        """The last run's build phases in ms (device wall-clock stamps at
        the build kernel's grid barriers; waits for the device)."""
        ms = np.zeros(len(BATCH_PHASES), np.float64)
        k = lib().rq_batch_phase_ms(self.ctx.h, self.h, _ptr(ms), ms.size)
        if k < 0:
            self.ctx._check(k)
        return dict(zip(BATCH_PHASES[:k], ms[:k].tolist()))

    def errors(self) -> int:
        bits = C.c_uint32()
        self.ctx._check(lib().rq_batch_errors(self.ctx.h, self.h, C.byref(bits)))
        return bits.value


class HostArray:
    """Page-locked host memory (rq_host_alloc) viewed as a numpy array."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        nbytes = max(self.count * self.dtype.itemsize, 1)
        self.ptr = lib().rq_host_alloc(nbytes)
        if not self.ptr:
            raise RQError(-2, f"rq_host_alloc({nbytes}) failed")
        buf = (C.c_char * nbytes).from_address(self.ptr)
        self.array = np.frombuffer(buf, self.dtype, self.count)

    def free(self) -> None:
        if self.ptr:
            self.array = None
            lib().rq_host_free(self.ptr)
            self.ptr = None


def _mesh_struct(m: MeshArrays):
    vx, vy, vz = (np.ascontiguousarray(a, np.float32) for a in (m.vx, m.vy, m.vz))
    idx = np.ascontiguousarray(m.idx, np.uint32)
    st = Mesh(vx.ctypes.data_as(C.POINTER(C.c_float)), vy.ctypes.data_as(C.POINTER(C.c_float)),
              vz.ctypes.data_as(C.POINTER(C.c_float)), vx.size,
              idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size // 3)
    return st, (vx, vy, vz, idx)


def server_clock() -> float:
    return lib().rq_server_clock()


class Server:
    """rq_server: resident mesh slots and a continuous-batching request loop.

    submit() keeps references to every array a request reads or writes until
    wait() / poll() reports it finished."""

    def __init__(self, device: int = 0, **cfg):
        c = ServerConfig()
        lib().rq_server_defaults(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.config = c
        h = C.c_void_p()
        rc = lib().rq_server_create(device, C.byref(c), C.byref(h))
        if rc != RQ_OK:
            raise (NoMemory if rc == -2 else RQError)(rc, f"rq_server_create failed ({rc})")
        self.h = h.value
        self._keep = {}

    def close(self) -> None:
        if self.h:
            lib().rq_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != RQ_OK:
            msg = lib().rq_server_last_error(self.h).decode()
            raise (NoMemory if rc == -2 else RQError)(rc, msg)

    def load_mesh(self, mesh: MeshArrays) -> int:
        st, keep = _mesh_struct(mesh)
        slot = C.c_uint32()
        self._check(lib().rq_server_load_mesh(self.h, C.byref(st), C.byref(slot)))
        return slot.value

    def submit(self, nrays: int, query: int, prim=None, t=None, occ=None, mesh: int = 0,
               new_mesh: MeshArrays | None = None, rays: dict | None = None,
               pinhole: dict | None = None, box=None, seed: int = 0) -> int:
        """prim / t / occ: numpy arrays (int32, float32, uint8) of >= nrays
        elements, ideally HostArray views. rays: host SoA dict (RAYS_HOST);
        pinhole: dict(eye, look_at, up, fov_deg, width); box: (lo, hi)."""
        r = Request()
        r.mesh = mesh
        r.nrays = nrays
        r.query = query
        keep = [prim, t, occ]
        if new_mesh is not None:
            st, arrays = _mesh_struct(new_mesh)
            keep += [st, arrays]
            r.new_mesh = C.pointer(st)
        if rays is not None:
            arrs = [np.ascontiguousarray(rays[f], np.float32) for f in RAY_FIELDS]
            keep.append(arrs)
            r.source = RAYS_HOST
            r.host_rays = Rays(*(a.ctypes.data_as(C.c_void_p) for a in arrs))
        elif pinhole is not None:
            r.source = RAYS_PINHOLE
            r.pinhole = Pinhole((C.c_float * 3)(*pinhole["eye"]),
                                (C.c_float * 3)(*pinhole["look_at"]),
                                (C.c_float * 3)(*pinhole.get("up", (0.0, 1.0, 0.0))),
                                pinhole.get("fov_deg", 40.0), pinhole.get("width", 1024))
        elif box is not None:
            r.source = RAYS_UNIFORM
            r.lo = (C.c_float * 3)(*box[0])
            r.hi = (C.c_float * 3)(*box[1])
        r.seed = seed
        r.prim = _ptr(prim)
        r.t = _ptr(t)
        r.occluded = _ptr(occ)
        tk = C.c_uint64()
        self._check(lib().rq_server_submit(self.h, C.byref(r), C.byref(tk)))
        self._keep[tk.value] = keep
        return tk.value

    def poll(self, ticket: int) -> RequestStatus:
        st = RequestStatus()
        self._check(lib().rq_server_poll(self.h, ticket, C.byref(st)))
        return st

    def wait(self, ticket: int) -> RequestStatus:
        st = RequestStatus()
        rc = lib().rq_server_wait(self.h, ticket, C.byref(st))
        if rc not in (RQ_OK,) and st.state >= 0:
            self._check(rc)
        return st

    def take_done(self, cap: int = 256) -> list[tuple[int, RequestStatus]]:
        """Requests finished since the last call: (ticket, status), completion order."""
        tk = (C.c_uint64 * cap)()
        st = (RequestStatus * cap)()
        n = C.c_uint32()
        self._check(lib().rq_server_take_done(self.h, tk, st, cap, C.byref(n)))
        return [(tk[i], st[i]) for i in range(n.value)]

    def forget(self, ticket: int) -> None:
        self._check(lib().rq_server_forget(self.h, ticket))
        self._keep.pop(ticket, None)

    def cancel(self, ticket: int) -> None:
        """rq_server_cancel: the request finishes with state RQ_E_CANCELED."""
        self._check(lib().rq_server_cancel(self.h, ticket))

    def stats(self) -> dict:
        st = ServerStats()
        self._check(lib().rq_server_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}


class Renderer:
    """rq_renderer: the GPU_LBVH wavefront replacement of render() (main.cpp:541-566)."""

    def __init__(self, ctx: Context, params: RenderParams):
        self.ctx = ctx
        self.params = params
        h = C.c_void_p()
        ctx._check(lib().rq_renderer_create(ctx.h, C.byref(params), C.byref(h)))
        self.h = h.value
        self.n = params.width * params.height * params.spp

    def frame(self, bvh: int, stream=None) -> None:
        self.ctx._check(lib().rq_render_frame(self.ctx.h, self.h, bvh, stream))

    def wait_traced(self, stream) -> None:
        """rq_render_wait_traced: `stream` waits (on the device) until the last
        frame has finished its last closest-hit launch."""
        self.ctx._check(lib().rq_render_wait_traced(self.ctx.h, self.h, stream))

    def counts(self) -> FrameCounts:
        c = FrameCounts()
        self.ctx._check(lib().rq_render_counts(self.ctx.h, self.h, C.byref(c)))
        return c

    def image(self) -> tuple[np.ndarray, np.ndarray]:
        npix = self.params.width * self.params.height
        rgb = np.empty(3 * npix, np.float32)
        ppm = np.empty(3 * npix, np.uint8)
        self.ctx._check(lib().rq_render_read_image(self.ctx.h, self.h, _ptr(rgb), _ptr(ppm)))
        return rgb, ppm

    def samples(self) -> dict:
        n = self.n
        out = dict(hit0=np.empty(n, np.int32), hit1=np.empty(n, np.int32),
                   shadow0=np.empty(n, np.uint8), shadow1=np.empty(n, np.uint8),
                   rgb=np.empty(3 * n, np.float32))
        self.ctx._check(lib().rq_render_read_samples(
            self.ctx.h, self.h, _ptr(out["hit0"]), _ptr(out["hit1"]), _ptr(out["shadow0"]),
            _ptr(out["shadow1"]), _ptr(out["rgb"])))
        return out

    def rays(self, which: int = 0) -> DeviceRays:
        """Non-owning view of a ray batch of the last frame (0 primary, 1 bounce, 2 shadow)."""
        r = Rays()
        n = C.c_uint32()
        self.ctx._check(lib().rq_render_rays(self.ctx.h, self.h, which, C.byref(r), C.byref(n)))
        view = DeviceRays.__new__(DeviceRays)
        view.n = n.value
        view.arrays = {}
        for f in RAY_FIELDS:
            a = DeviceArray.__new__(DeviceArray)
            a.ctx, a.nbytes, a.ptr, a.owned = self.ctx, 4 * n.value, getattr(r, f), False
            view.arrays[f] = a
        return view

    def primary_rays(self) -> DeviceRays:
        return self.rays(0)

    def close(self) -> None:
        if self.h:
            lib().rq_renderer_destroy(self.ctx.h, self.h)
            self.h = None


class FramePipeline:
    """Frames that rebuild their LBVH every frame, the rebuild for frame k+1
    overlapping the render of frame k on the device.

    Two BVHs of the scene alternate, each rebuilt on its own stream
    (rq_rebuild_lbvh, then rq_bvh_treelets when `treelets` > 0); the renderer
    runs on a third. rq_stream_order makes a render wait for its own BVH's
    rebuild and a rebuild wait for the render that last read its BVH, with no
    host wait. `overlap` "frame": the rebuild for frame k+1 is ordered after
    the render of frame k-1 and runs beside all of frame k (for short frames
    whose launches leave the device idle); "tail": it is ordered after frame
    k's last closest-hit launch (rq_render_wait_traced) and runs beside the
    frame's shading and second-hit shadow rays only, so it does not compete
    with a traversal that keeps every SIMD issuing. Not pipelined
    (`pipelined` False): rebuild and render one after the other on the
    default stream, with one BVH. The scene must not change while frames are
    in flight. frame() returns the BVH handle it rendered.
    """

    def __init__(self, ctx: Context, renderer: Renderer, scene: int, bvh: int,
                 morton_mode: int = MORTON_SCENE_AABB, treelets: int = 0,
                 pipelined: bool = True, overlap: str = "frame"):
        if overlap not in ("frame", "tail"):
            raise ValueError(f"overlap must be 'frame' or 'tail', not {overlap!r}")
        self.ctx, self.ren, self.treelets, self.pipelined = ctx, renderer, treelets, pipelined
        self.overlap = overlap
        self.bvhs = [bvh]
        self.streams: list = []
        self.k = 0
        if pipelined:
            self.bvhs.append(ctx.build_lbvh(scene, morton_mode))
            self.streams = [ctx.stream_create() for _ in range(3)]

    def _rebuild(self, b: int, stream) -> None:
        self.ctx.rebuild(b, stream=stream)
        if self.treelets:
            self.ctx.treelets(b, self.treelets, stream=stream)

    def frame(self) -> int:
        if not self.pipelined:
            self._rebuild(self.bvhs[0], None)
            self.ren.frame(self.bvhs[0])
            return self.bvhs[0]
        k, s_build, s_render = self.k, self.streams[:2], self.streams[2]
        cur, nxt = k % 2, (k + 1) % 2
        if k == 0:
            self._rebuild(self.bvhs[cur], s_build[cur])
        if self.overlap == "frame":
            self.ctx.stream_order(s_build[nxt], s_render)  # after the render that read bvhs[nxt]
            self._rebuild(self.bvhs[nxt], s_build[nxt])    # frame k+1's BVH
            self.ctx.stream_order(s_render, s_build[cur])  # after this frame's rebuild
            self.ren.frame(self.bvhs[cur], stream=s_render)
        else:
            self.ctx.stream_order(s_render, s_build[cur])  # after this frame's rebuild
            self.ren.frame(self.bvhs[cur], stream=s_render)
            # frame k+1's BVH once frame k has traced: frame k-1, the last
            # reader of bvhs[nxt], ran before it on the same stream
            self.ren.wait_traced(s_build[nxt])
            self._rebuild(self.bvhs[nxt], s_build[nxt])
        self.k = k + 1
        return self.bvhs[cur]

    def render_stream(self):
        """The stream frames render on (None when not pipelined)."""
        return self.streams[2] if self.pipelined else None

    def close(self) -> None:
        """Waits for the frames in flight, frees the second BVH and the streams
        (the BVH passed in stays the caller's)."""
        self.ctx.sync()
        for s in self.streams:
            self.ctx.stream_destroy(s)
        self.streams = []
        for b in self.bvhs[1:]:
            self.ctx.release(b)
        self.bvhs = self.bvhs[:1]


def write_ppm(path: str | Path, w: int, h: int, data: np.ndarray) -> None:
    data = np.ascontiguousarray(data, np.uint8)
    rc = lib().rq_write_ppm(str(path).encode(), w, h, _ptr(data))
    if rc != RQ_OK:
        raise RQError(rc, f"write_ppm {path}")


def bvh_blob_validate(blob: bytes, prims) -> tuple[int, str]:
    """rq_bvh_blob_validate (host only, no device): (status, reason) of a BVH
    blob against a triangle mesh (MeshArrays) or a sphere set (SphereArrays)."""
    b = np.frombuffer(blob, np.uint8)
    msg = C.create_string_buffer(256)
    if hasattr(prims, "cx"):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (prims.cx, prims.cy, prims.cz, prims.r)]
        rc = lib().rq_bvh_blob_validate(_ptr(b), b.size, PRIM_SPHERE, _ptr(arrs[0]), _ptr(arrs[1]),
                                        _ptr(arrs[2]), arrs[0].size, None, _ptr(arrs[3]),
                                        arrs[0].size, msg, 256)
    else:
        arrs = [np.ascontiguousarray(a, np.float32) for a in (prims.vx, prims.vy, prims.vz)]
        idx = np.ascontiguousarray(prims.idx, np.uint32)
        rc = lib().rq_bvh_blob_validate(_ptr(b), b.size, PRIM_TRIANGLE, _ptr(arrs[0]),
                                        _ptr(arrs[1]), _ptr(arrs[2]), arrs[0].size, _ptr(idx),
                                        None, idx.size // 3, msg, 256)
    return rc, msg.value.decode()
