"""A compiled C++ program (tests/abi/caller.cpp) drives librq through the C
ABI as a reference-side GPU_LBVH case would (OBJ read, upload, rq_build_lbvh,
then device-batch, host-batch and per-ray queries); its results must equal
the oracle's: hit ids and t bit for bit, occlusion flags identical
(main.cpp:343-358 and :373-387, lowest index on ties). The Python host-pointer
bindings are checked the same way."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import log
from helpers import RAY_FIELDS, concat_rays, edge_rays, random_rays, rays_toward

pytestmark = pytest.mark.gpu

CALLER = Path(__file__).resolve().parent / "abi" / "caller"


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _write_rays(path, rays):
    n = len(rays["ox"])
    with open(path, "wb") as f:
        f.write(np.uint32(n).tobytes())
        for k in RAY_FIELDS:
            f.write(np.ascontiguousarray(rays[k], np.float32).tobytes())


def _read_out(path):
    b = Path(path).read_bytes()
    n, k = np.frombuffer(b[:8], np.uint32)
    n, k = int(n), int(k)
    off = 8
    out = {}

    def take(name, dtype, count):
        nonlocal off
        a = np.frombuffer(b, dtype, count, off)
        off += a.nbytes
        out[name] = a

    for tag, m in (("dev", n), ("host", n), ("ray", k)):
        take(f"{tag}_prim", np.int32, m)
        take(f"{tag}_t", np.float32, m)
        take(f"{tag}_occ", np.uint8, m)
    take("info", np.uint32, 2)
    assert off == len(b)
    return n, k, out


def _rays():
    return concat_rays(rays_toward(6000, 41), random_rays(3000, 42), edge_rays())


def test_compiled_caller_matches_the_oracle(rq, O, tmp_path):
    assert CALLER.exists(), "tests/abi/caller is not built (tests/abi/build_caller.py)"
    mesh0 = rq.icosphere(5, 7, center=(0, 0, -5))
    obj = tmp_path / "mesh.obj"
    rq.obj_write(obj, mesh0)
    mesh = rq.obj_read(obj)  # the mesh exactly as the program reads it
    rays = _rays()
    _write_rays(tmp_path / "rays.bin", rays)
    log(f"caller: {mesh.ntri} triangles, {len(rays['ox'])} rays")
    r = subprocess.run([str(CALLER), str(obj), str(tmp_path / "rays.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    log(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    n, k, g = _read_out(tmp_path / "out.bin")
    assert n == len(rays["ox"]) and k == min(n, 1000)
    ob = O.OracleBVH(mesh, rq.MORTON_SCENE_AABB)
    o, oa = ob.closest(rays), ob.any(rays)
    assert g["info"][0] == mesh.ntri and g["info"][1] == ob.height
    assert (o["prim"] >= 0).sum() > 1000 and (oa["occ"] == 0).sum() > 100
    for tag in ("dev", "host"):
        assert np.array_equal(g[f"{tag}_prim"], o["prim"]), tag
        assert np.array_equal(_bits(g[f"{tag}_t"]), _bits(o["t"])), tag
        assert np.array_equal(g[f"{tag}_occ"], oa["occ"]), tag
    assert np.array_equal(g["ray_prim"], o["prim"][:k])
    assert np.array_equal(_bits(g["ray_t"]), _bits(o["t"][:k]))
    assert np.array_equal(g["ray_occ"], oa["occ"][:k])


def test_host_pointer_queries(ctx, rq, O):
    """rq_closest_hit_host / rq_any_hit_host (staged in launches of 2^20 rays:
    2^20 + 4097 rays take two) and the per-ray rq_cast_ray / rq_occluded_ray."""
    mesh = rq.icosphere(4, 3, center=(0, 0, -5))
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    try:
        ob = O.OracleBVH(mesh, rq.MORTON_SCENE_AABB)
        for rays in (_rays(), rays_toward((1 << 20) + 4097, 43)):
            n = len(rays["ox"])
            log(f"host queries: {n} rays")
            g = ctx.closest_hit_host(b, rays)
            dev = ctx.closest_hit(b, rays)
            for key in ("prim", "t", "u", "v"):
                assert np.array_equal(_bits(g[key]) if key != "prim" else g[key],
                                      _bits(dev[key]) if key != "prim" else dev[key]), key
            occ = ctx.any_hit_host(b, rays)
            assert np.array_equal(occ, ctx.any_hit(b, rays)["occ"])
            sl = slice(0, n, max(1, n // 20000))
            sub = {k: v[sl] for k, v in rays.items()}
            o = ob.closest(sub)
            assert np.array_equal(g["prim"][sl], o["prim"])
            assert np.array_equal(_bits(g["t"][sl]), _bits(o["t"]))
        rays = _rays()
        o, oa = ob.closest(rays), ob.any(rays)
        for i in list(range(0, 9000, 450)) + list(range(9000, len(rays["ox"]), 7)):
            orig = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
            d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
            prim, t, uv = ctx.cast_ray(b, orig, d, rays["tmin"][i], rays["tmax"][i])
            assert prim == o["prim"][i], i
            assert np.float32(t).view(np.uint32) == o["t"][i:i + 1].view(np.uint32)[0], i
            if prim >= 0:
                assert np.float32(uv[0]).view(np.uint32) == _bits(o["u"][i:i + 1])[0], i
            assert ctx.occluded_ray(b, orig, d, rays["tmin"][i], rays["tmax"][i]) == bool(
                oa["occ"][i]), i
        assert ctx.device_errors() == 0
    finally:
        ctx.release(b)
        ctx.release(s)
