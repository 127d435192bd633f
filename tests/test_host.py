"""CPU: the C-ABI library loads and exports every declared entry point; the
host-side scene generators and OBJ I/O; the dataStructure switch value."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import REFERENCE, ROOT

HEADERS = sorted((ROOT / "include" / "rq").glob("*.h"))


def _declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(rq_\w+)\s*\(", text, re.M):
            if "static inline" not in text[max(0, m.start() - 40):m.start() + 20]:
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(rq):
    lib = rq.lib()
    names = _declared()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "rq/rq.h"\n#include "rq/rq_scene.h"\n#include "rq/rq_render.h"\n'
                   '#include "rq/settings.h"\n#include "rq/rq_batch.h"\n#include "rq/rq_serve.h"\n'
                   "_Static_assert(RQ_BVH == 0 && RQ_KDTREE == 1 && RQ_UNIFORM_GRID == 2 && "
                   "RQ_LBVH == 3 && RQ_NONE == 4 && RQ_GPU_LBVH == 5, \"AccType\");\n"
                   "int main(void) { rq_settings s = rq_settings_default();\n"
                   "  return s.data_structure == RQ_GPU_LBVH ? 0 : 1; }\n")
    inc = str(ROOT / "include")
    subprocess.run(["gcc", "-std=c11", "-fsyntax-only", f"-I{inc}", str(src)], check=True)
    cpp = tmp_path / "t.cpp"
    cpp.write_text(src.read_text().replace("_Static_assert", "static_assert"))
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{inc}", str(cpp)], check=True)


def test_trace_variant_table(rq):
    """The kernel variant table (a static table, no device needed): variant 0
    is the canonical binary walk; every wide any-hit kernel whose LDS stack
    leaves room for six waves per SIMD (6 x 4 waves x 64 lanes x depth x 4 B
    <= 160 KB) is capped for six; 17 and 18 cap their closest-hit kernels too."""
    vs = rq.Context.trace_variants()
    assert vs[0] == dict(index=0, block=128, lds_depth=24, tstack=0, wide=0, closest_waves=0,
                         any_waves=0)
    for v in vs:
        fits6 = v["wide"] and 6 * 4 * 64 * v["lds_depth"] * 4 <= 160 * 1024
        assert v["any_waves"] == (6 if fits6 else 0), v
        assert v["closest_waves"] in (0, 6)
    six = [v["index"] for v in vs if v["closest_waves"] == 6]
    assert six == [17, 18] and all(vs[i]["wide"] and not vs[i]["tstack"] for i in six)


def test_acc_type_values(rq):
    assert [int(x) for x in rq.AccType] == [0, 1, 2, 3, 4, 5]
    assert rq.AccType.GPU_LBVH.name == "GPU_LBVH"


def test_context_fails_loudly_without_gpu(rq):
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(rq.RQError):
        rq.Context(0)


@pytest.mark.parametrize("level", [0, 1, 2, 3, 5])
def test_icosphere_counts_and_watertight(rq, level):
    m = rq.icosphere(level, 1)
    assert m.nv == 10 * 4 ** level + 2 and m.ntri == 20 * 4 ** level
    f = m.idx.reshape(-1, 3).astype(np.int64)
    directed = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    key = directed[:, 0] * m.nv + directed[:, 1]
    assert len(np.unique(key)) == len(key)            # consistent orientation
    rev = directed[:, 1] * m.nv + directed[:, 0]
    assert np.isin(rev, key).all()                     # every edge shared by two faces
    r = np.sqrt(m.vx.astype(np.float64) ** 2 + m.vy ** 2 + m.vz ** 2)
    assert r.min() >= 0.9 - 1e-5 and r.max() <= 1.1 + 1e-5


def test_generators_deterministic(rq):
    a, b, c = rq.icosphere(4, 7), rq.icosphere(4, 7), rq.icosphere(4, 8)
    assert np.array_equal(a.vx, b.vx) and np.array_equal(a.idx, b.idx)
    assert not np.array_equal(a.vx, c.vx) and np.array_equal(a.idx, c.idx)
    s1, s2 = rq.soup(1000, 3), rq.soup(1000, 3)
    assert np.array_equal(s1.vy, s2.vy)


def test_icosphere_field(rq):
    m = rq.icosphere_field(64, 1, 3, 3, 100.0, 2.5)
    assert m.ntri == 64 * 80 and m.nv == 64 * 42
    v = np.stack([m.vx, m.vy, m.vz], 1).reshape(64, 42, 3)
    centres = v.mean(axis=1)
    d = np.linalg.norm(centres[:, None] - centres[None], axis=2) + np.eye(64) * 1e9
    assert d.min() > 2.0
    assert (m.idx.reshape(64, -1) // 42 == np.arange(64)[:, None]).all()


def test_soup_edges(rq):
    m = rq.soup(20000, 5, 0.02, 0.005, 0.002)
    v = np.stack([m.vx, m.vy, m.vz], 1).astype(np.float64)[m.idx].reshape(-1, 3, 3)
    e = np.linalg.norm(v[:, 1] - v[:, 0], axis=1)
    assert abs(e.mean() - 0.02) < 5e-4 and e.min() >= 0.002 - 1e-6
    c = v.mean(axis=1)
    assert c.min() >= -0.05 and c.max() <= 1.05


def test_obj_round_trip(rq, tmp_path):
    m = rq.icosphere(3, 2, center=(1.5, -2.0, -7.0))
    p = tmp_path / "m.obj"
    rq.obj_write(p, m)
    r = rq.obj_read(p)
    assert np.array_equal(r.vx, m.vx) and np.array_equal(r.vz, m.vz)
    assert np.array_equal(r.idx, m.idx)


def test_obj_forms(rq, tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("# comment\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
                 "f 1/1/1 2/2/1 3/3/1 4/4/1\nf -4 -3 -1\n")
    m = rq.obj_read(p)
    assert m.nv == 4 and m.idx.tolist() == [0, 1, 2, 0, 2, 3, 0, 1, 3]
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(rq.RQError):
        rq.obj_read(bad)


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference models not present")
def test_reference_models_vertex_counts(rq):
    """The reference's models are vertex-only point clouds (createScene_new
    reads `v` lines only, main.cpp:668-697)."""
    for name, nv in [("bunny.obj", 35947), ("teapot.obj", 3644), ("woody.obj", 694)]:
        m = rq.obj_read(REFERENCE / "models" / name)
        assert m.nv == nv and m.ntri == 0


def test_render_defaults_are_reference_constants(rq):
    p = rq.render_defaults()
    assert (p.width, p.height, p.spp, p.bounce) == (640, 480, 1, 0)
    assert list(p.light) == [0.0, 3.0, 30.0]          # light2, main.cpp:775
    assert np.allclose(list(p.sky), [0.6, 0.8, 1.0])   # main.cpp:312
    assert p.bias == np.float32(1e-4)                  # main.cpp:404


def test_frame_pipeline_rejects_unknown_overlap(rq):
    """rq.FramePipeline takes overlap "frame" or "tail" and rejects anything
    else before it touches the device."""
    with pytest.raises(ValueError):
        rq.FramePipeline(None, None, 0, 0, overlap="tails")
