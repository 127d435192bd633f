"""GPU LBVH build parity: Morton codes, sorted order, Karras topology and
refit boxes bit-exact with the CPU oracle (oracle/oracle.cpp)."""
import numpy as np
import pytest

from conftest import log
from helpers import duplicate_mesh, mesh_configs, random_soup

pytestmark = pytest.mark.gpu


def _check_build(ctx, O, rq, mesh, mode):
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, mode)
    try:
        g = ctx.export(b, mesh.ntri)
        ob = O.OracleBVH(mesh, mode)
        o = ob.arrays()
        assert np.array_equal(g["codes"], o["codes"]), "morton codes"
        assert np.array_equal(g["sorted_idx"], o["sorted_idx"]), "sorted order"
        assert np.array_equal(g["left"], o["left"]), "left children"
        assert np.array_equal(g["right"], o["right"]), "right children"
        assert np.array_equal(g["boxes"].view(np.uint32), o["boxes"].view(np.uint32)), "boxes"
        info = ctx.bvh_info(b)
        assert info.height == ob.height
        assert info.n_prims == mesh.ntri
    finally:
        ctx.release(b)
        ctx.release(s)


@pytest.mark.parametrize("name", list(mesh_configs()))
@pytest.mark.parametrize("mode", [0, 1])
def test_build_small(ctx, O, rq, name, mode):
    log(f"build {name} mode={mode}")
    _check_build(ctx, O, rq, mesh_configs()[name](), mode)


@pytest.mark.parametrize("n", [2, 3, 4095, 4096, 4097, 8192 + 17, 65537])
def test_build_ragged_soup(ctx, O, rq, n):
    log(f"build soup n={n}")
    _check_build(ctx, O, rq, random_soup(n, n, spread=2.0, size=0.01), 1)


def test_build_all_duplicates(ctx, O, rq):
    log("build duplicates 5000")
    _check_build(ctx, O, rq, duplicate_mesh(5000), 1)


def test_build_c1_mesh(ctx, O, rq):
    """81,920-triangle displaced icosphere (subdivision 6), both Morton maps."""
    m = rq.icosphere(6, 1, center=(0, 0, -5))
    assert m.ntri == 81920 and m.nv == 40962
    for mode in (0, 1):
        log(f"build c1 mode={mode}")
        _check_build(ctx, O, rq, m, mode)


@pytest.mark.slow
def test_build_c2_mesh_and_rebuild(ctx, O, rq):
    """1,310,720 triangles (subdivision 8): build, then rebuild in place
    twice (the per-frame rebuild) and compare every time."""
    m = rq.icosphere(8, 2, center=(0, 0, -2))
    assert m.ntri == 1310720
    s = ctx.upload_scene(m)
    b = ctx.build_lbvh(s, 1)
    try:
        ob = O.OracleBVH(m, 1)
        o = ob.arrays()
        for it in range(3):
            log(f"build c2 iteration {it}")
            if it:
                ctx.rebuild(b)
            g = ctx.export(b, m.ntri)
            for k in ("codes", "sorted_idx", "left", "right"):
                assert np.array_equal(g[k], o[k]), k
            assert np.array_equal(g["boxes"].view(np.uint32), o["boxes"].view(np.uint32))
    finally:
        ctx.release(b)
        ctx.release(s)


def test_empty_scene(ctx, rq):
    log("empty scene")
    m = rq.MeshArrays(np.zeros(0, np.float32), np.zeros(0, np.float32),
                      np.zeros(0, np.float32), np.zeros(0, np.uint32))
    s = ctx.upload_scene(m)
    b = ctx.build_lbvh(s, 1)
    info = ctx.bvh_info(b)
    assert info.n_prims == 0
    ctx.release(b)
    ctx.release(s)
