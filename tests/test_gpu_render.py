"""GPU wavefront renderer parity with the CPU oracle: camera rays bit-exact,
primary/bounce hit ids and shadow flags identical, image within 1/255 per
channel (the only non-correctly-rounded operation is powf in the specular
term, main.cpp:475)."""
import numpy as np
import pytest

from conftest import log

pytestmark = pytest.mark.gpu

RAYS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def _render_parity(ctx, O, rq, mesh, params, sub=1):
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    r = rq.Renderer(ctx, params)
    try:
        r.frame(b)
        ctx.sync()
        assert ctx.device_errors() == 0
        n = params.width * params.height * params.spp
        g = r.samples()
        rgb, ppm = r.image()
        cnt = r.counts()
        prim = r.rays(0).download()
        if params.camera == rq.CAMERA_REF_JITTER:
            orays = O.camera_rays_ref(params.width, params.height, params.spp,
                                      O.mt19937_jitter(2 * n))
        else:
            orays = O.camera_rays_philox(params.width, params.height, params.spp, params.seed)
        for k in RAYS:
            assert np.array_equal(prim[k].view(np.uint32), orays[k].view(np.uint32)), k
        ob = O.OracleBVH(mesh, rq.MORTON_SCENE_AABB)
        sp = O.shade_params(light=tuple(params.light), emission=tuple(params.emission),
                            surface=tuple(params.surface), sky=tuple(params.sky),
                            bias=params.bias, bounce=params.bounce)
        o = O.render_samples(ob, mesh, orays, sp)
        assert np.array_equal(g["hit0"], o["hit0"]), "primary hit ids"
        assert np.array_equal(g["shadow0"], o["shadow0"]), "primary shadow flags"
        if params.bounce:
            assert np.array_equal(g["hit1"], o["hit1"]), "bounce hit ids"
            assert np.array_equal(g["shadow1"], o["shadow1"]), "bounce shadow flags"
        # ray counts: oracle counts[0] = all traced rays
        assert cnt.primary + cnt.bounce + cnt.shadow == int(o["counts"][0])
        # colour: per sample within 2e-5 absolute (powf differs in the last ulps)
        assert np.max(np.abs(g["rgb"] - o["rgb"])) < 2e-5
        oppm = O.resolve_ppm(o["rgb"], params.width * params.height, params.spp)
        diff = np.abs(ppm.astype(np.int16) - oppm.astype(np.int16))
        assert diff.max() <= 1, f"image differs by {diff.max()}/255"
        return g, ppm, cnt
    finally:
        r.close()
        ctx.release(b)
        ctx.release(s)


def test_render_c1_reference_camera(ctx, O, rq):
    """Config 1: 81,920-triangle displaced icosphere, 640x480, 1 spp with the
    reference's mt19937 jitter, one light, Lambert + one shadow ray per hit."""
    mesh = rq.icosphere(6, 1, center=(0.0, 0.0, -5.0))
    p = rq.render_defaults()
    log("render c1")
    g, ppm, cnt = _render_parity(ctx, O, rq, mesh, p)
    hits = (g["hit0"] >= 0).sum()
    assert cnt.primary == 640 * 480 and cnt.shadow == hits and cnt.bounce == 0
    assert hits > 0.2 * 640 * 480
    # sky pixels carry the reference's miss colour (0.6, 0.8, 1) -> (153, 204, 255)
    sky = ppm.reshape(-1, 3)[g["hit0"] < 0]
    assert (sky == np.array([153, 204, 255], np.uint8)).all()


def test_render_bounce_philox(ctx, O, rq):
    """Config 2 shape at reduced size: 4 spp Philox jitter, mirror bounce and
    shadow rays from both hits, mesh close to the camera."""
    mesh = rq.icosphere(6, 2, center=(0.0, 0.0, -2.0))
    p = rq.render_defaults()
    p.width, p.height, p.spp = 320, 180, 4
    p.camera, p.seed, p.bounce = rq.CAMERA_PHILOX, 2, 1
    log("render bounce 320x180x4")
    g, ppm, cnt = _render_parity(ctx, O, rq, mesh, p)
    assert cnt.bounce == (g["hit0"] >= 0).sum()
    assert cnt.shadow == (g["hit0"] >= 0).sum() + (g["hit1"] >= 0).sum()
    assert (g["hit1"] >= 0).sum() > 0


def test_render_surface_colour(ctx, O, rq):
    """Non-zero surfaceColor (added per light, main.cpp:490) and a light
    placed so that parts of the mesh are in shadow."""
    mesh = rq.icosphere(5, 7, amp=0.3, center=(0.0, 0.0, -4.0))
    p = rq.render_defaults()
    p.width, p.height = 200, 150
    p.surface[0], p.surface[1], p.surface[2] = 0.1, 0.05, 0.2
    p.light[0], p.light[1], p.light[2] = 3.0, 4.0, -2.0
    log("render surface colour")
    g, ppm, cnt = _render_parity(ctx, O, rq, mesh, p)
    assert g["shadow0"].sum() > 0


def test_render_frames_back_to_back_on_a_stream(ctx, rq):
    """Several bounce + shadow frames queued on a caller stream without waits
    in between (the primary shadow rays run on the renderer's second stream
    and are joined back by events): every frame equals a single frame."""
    mesh = rq.icosphere(5, 3, center=(0.0, 0.0, -2.5))
    p = rq.render_defaults()
    p.width, p.height, p.spp = 160, 96, 2
    p.camera, p.seed, p.bounce = rq.CAMERA_PHILOX, 5, 1
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    r = rq.Renderer(ctx, p)
    st = ctx.stream_create()
    try:
        r.frame(b)
        ctx.sync()
        ref = r.samples()
        ref_rgb, ref_ppm = r.image()
        for _ in range(3):
            r.frame(b, stream=st)
        ctx.sync(st)
        g = r.samples()
        for k in ("hit0", "hit1", "shadow0", "shadow1"):
            assert np.array_equal(g[k], ref[k]), k
        assert np.array_equal(g["rgb"].view(np.uint32), ref["rgb"].view(np.uint32))
        rgb, ppm = r.image()
        assert np.array_equal(ppm, ref_ppm)
        assert ctx.device_errors() == 0
    finally:
        ctx.stream_destroy(st)
        r.close()
        ctx.release(b)
        ctx.release(s)


@pytest.mark.parametrize("overlap,bounce,occlusion",
                         [("frame", 1, False), ("tail", 1, False), ("tail", 0, False),
                          ("tail", 0, True)])
def test_pipelined_frames_equal_single_frames(ctx, rq, overlap, bounce, occlusion):
    """rq.FramePipeline: the LBVH rebuild for frame k+1 runs on its own stream
    beside frame k (two BVHs alternating, ordered by rq_stream_order, or, with
    overlap "tail", after frame k's last closest-hit launch by
    rq_render_wait_traced); every frame equals a frame rendered after a plain
    rebuild, bit for bit."""
    from rq import workloads as W
    mesh, p, _ = W.c1()
    p.bounce = bounce
    if occlusion:  # the traced event follows the frame's any-hit launch
        p.mode = rq.RENDER_OCCLUSION
    s = ctx.upload_scene(mesh)
    b = ctx.build_lbvh(s, rq.MORTON_SCENE_AABB)
    r = rq.Renderer(ctx, p)
    pipe = None
    try:
        ctx.rebuild(b)
        r.frame(b)
        ctx.sync()
        ref = None if occlusion else r.samples()
        ref_rgb, ref_ppm = r.image()
        pipe = rq.FramePipeline(ctx, r, s, b, rq.MORTON_SCENE_AABB, overlap=overlap)
        used = []
        for k in range(8):
            log(f"pipelined frame {k} overlap {overlap} bounce {bounce}")
            used.append(pipe.frame())
            if k % 2 == 1:  # frames 0, 2, ... are left in flight behind the next one
                ctx.sync(pipe.render_stream())
                if ref is not None:
                    g = r.samples()
                    for key in ("hit0", "hit1", "shadow0", "shadow1"):
                        assert np.array_equal(g[key], ref[key]), (k, key)
                    assert np.array_equal(g["rgb"].view(np.uint32), ref["rgb"].view(np.uint32)), k
                rgb, ppm = r.image()
                assert np.array_equal(rgb.view(np.uint32), ref_rgb.view(np.uint32)), k
                assert np.array_equal(ppm, ref_ppm), k
        assert used[0] == b and used[1] != b and used[2] == b  # the two BVHs alternate
        assert ctx.device_errors() == 0
    finally:
        if pipe is not None:
            pipe.close()
        r.close()
        ctx.release(b)
        ctx.release(s)


def test_stream_order(ctx, rq):
    """rq_stream_order: a copy enqueued on one stream after a fill on another
    sees the fill, with no host wait in between."""
    nbytes = 256 << 20
    a = rq.DeviceArray(ctx, nbytes)
    out = rq.DeviceArray(ctx, 4096)
    s1, s2 = ctx.stream_create(), ctx.stream_create()
    try:
        lib = rq.lib()
        for value in (7, 9, 11):
            a.fill(0)
            out.fill(0)
            ctx.sync()
            a.fill(value, stream=s1)
            ctx.stream_order(s2, s1)
            assert lib.rq_memcpy_d2d(ctx.h, out.ptr, a.ptr + nbytes - 4096, 4096, s2) == 0
            ctx.sync(s2)
            assert (out.download(np.uint8, 4096, stream=s2) == value).all(), value
        ctx.stream_order(s1, s1)  # one stream: nothing to order
    finally:
        ctx.stream_destroy(s1)
        ctx.stream_destroy(s2)
        a.free()
        out.free()
