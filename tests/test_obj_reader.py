"""CPU: the parallel memory-mapped OBJ reader gives the same mesh for every
thread count and equals an independent restatement of its rules (v records,
f records with 1-based / negative indices, a/b/c tokens and polygon fans;
every other line ignored)."""
import numpy as np
import pytest


def _py_parse(text):
    """Independent sequential restatement of the reader's rules."""
    xs, idx = [], []
    for line in text.split("\n"):
        s = line.lstrip(" \t")
        if len(s) > 1 and s[0] == "v" and s[1] in " \t":
            xs.append([np.float32(float(t)) for t in s[1:].split()[:3]])
        elif len(s) > 1 and s[0] == "f" and s[1] in " \t":
            poly = []
            for tok in s[1:].replace("\r", " ").split():
                v = int(tok.split("/")[0])
                i0 = v - 1 if v > 0 else len(xs) + v
                assert v != 0 and 0 <= i0 < len(xs)
                poly.append(i0)
            for k in range(1, len(poly) - 1):
                idx += [poly[0], poly[k], poly[k + 1]]
    v = np.array(xs, np.float32).reshape(-1, 3)
    return v, np.array(idx, np.uint32)


def _random_obj(rng, nv_blocks=40):
    lines, nv = ["# header comment", "o thing", ""], 0
    for _ in range(nv_blocks):
        for _ in range(int(rng.integers(3, 60))):
            x, y, z = rng.normal(size=3)
            sep = " " if rng.random() < 0.8 else "\t"
            lines.append(f"v{sep}{x:.7g} {y:.7g} {z:.7g}" + ("\r" if rng.random() < 0.1 else ""))
            nv += 1
        lines.append("vn 0 0 1")
        for _ in range(int(rng.integers(1, 40))):
            k = int(rng.integers(3, 7))
            toks = []
            for _ in range(k):
                i = int(rng.integers(0, nv))
                tok = str(i + 1) if rng.random() < 0.6 else str(i - nv)
                form = rng.random()
                if form < 0.2:
                    tok += "/1/1"
                elif form < 0.3:
                    tok += "//2"
                toks.append(tok)
            lines.append("  f " + " ".join(toks))
        lines.append("s off")
    return "\n".join(lines)  # no trailing newline


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_obj_reader_parallel_equals_sequential_and_rules(rq, tmp_path, seed):
    text = _random_obj(np.random.default_rng(seed))
    p = tmp_path / f"r{seed}.obj"
    p.write_text(text)
    v, idx = _py_parse(text)
    for threads in (1, 2, 7, 16):
        m = rq.obj_read(p, threads)
        assert np.array_equal(np.stack([m.vx, m.vy, m.vz], 1), v), threads
        assert np.array_equal(m.idx, idx), threads


def test_obj_reader_large_roundtrip(rq, tmp_path):
    mesh = rq.icosphere(6, 4)
    p = tmp_path / "ico.obj"
    rq.obj_write(p, mesh)
    a, b = rq.obj_read(p, 1), rq.obj_read(p, 16)
    for k in ("vx", "vy", "vz", "idx"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
        assert np.array_equal(getattr(a, k), getattr(mesh, k))


@pytest.mark.parametrize("bad", ["v 1 2 3\nf 1 2 4\n", "v 1 2 3\nf 0 1 1\n",
                                 "v 1 2 3\nv 1 2 3\nv 1 2 3\nf 1 2 -4\n", "v 1 2 3\nf 1 x 2\n"])
def test_obj_reader_rejects_bad_records(rq, tmp_path, bad):
    p = tmp_path / "bad.obj"
    p.write_text(bad)
    for threads in (1, 4):
        with pytest.raises(Exception):
            rq.obj_read(p, threads)


def test_obj_reader_missing_coordinate_reads_zero(rq, tmp_path):
    p = tmp_path / "short.obj"
    p.write_text("v 1 2\nv 3 4 5\n")
    for threads in (1, 2):
        m = rq.obj_read(p, threads)
        assert np.array_equal(np.stack([m.vx, m.vy, m.vz], 1), np.float32([[1, 2, 0], [3, 4, 5]]))
