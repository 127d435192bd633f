"""The ctypes mirrors of the C ABI structs (rq/__init__.py) against the
headers: a C program compiled from include/rq/*.h prints every field's offset
and every struct's size, which must equal ctypes' (a field renamed in a
header fails to compile, a field added, moved or retyped changes an offset)."""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

PAIRS = [("Rays", "rq_rays"), ("Hits", "rq_hits"), ("BvhInfo", "rq_bvh_info"),
         ("TraceVariant", "rq_trace_variant"), ("Mesh", "rq_mesh"),
         ("RenderParams", "rq_render_params"), ("Spheres", "rq_spheres"),
         ("FrameCounts", "rq_frame_counts"), ("BatchItem", "rq_batch_item"),
         ("BatchConfig", "rq_batch_config"), ("Pinhole", "rq_pinhole"),
         ("Request", "rq_request"), ("RequestStatus", "rq_request_status"),
         ("ServerConfig", "rq_server_config"), ("ServerStats", "rq_server_stats")]


def test_ctypes_mirrors_match_the_headers(rq, tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "rq/rq.h"',
             '#include "rq/rq_batch.h"', '#include "rq/rq_render.h"', '#include "rq/rq_scene.h"',
             '#include "rq/rq_serve.h"', "int main(void) {"]
    expect = []
    for py, cname in PAIRS:
        cls = getattr(rq, py)
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(cls))
        for f, _ in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {f}));')
            expect.append(getattr(cls, f).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert len(got) == len(expect)
    i = 0
    for py, cname in PAIRS:
        cls = getattr(rq, py)
        assert got[i] == expect[i], f"sizeof({cname}) {got[i]} != ctypes {expect[i]}"
        i += 1
        for f, _ in cls._fields_:
            assert got[i] == expect[i], f"{cname}.{f}: offset {got[i]} != ctypes {expect[i]}"
            i += 1
