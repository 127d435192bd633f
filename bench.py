#!/usr/bin/env python3
"""Benchmark of the GPU_LBVH ray-query path on one MI355X.

Default workload (incoherent large mesh): displaced icosphere, subdivision 8
(1,310,720 FP32 triangles) centred at (0,0,-2); 1920x1080, 4 jittered samples
per pixel (Philox); one mirror bounce per primary hit plus a shadow ray from
every hit; the 30-bit Morton LBVH is rebuilt every frame. One step = one
frame: LBVH rebuild + raygen + closest hit (primary) + shade + closest hit
(bounce) + shade + any hit (shadow) + resolve, inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--config c2|c1|c3|c4|c5]

Prints one JSON line: Mrays/s over all rays traced per frame; the roofline of
the dominant kernel and of the closest-hit kernel (algorithmic bytes per
launch / HIP-event launch time against 8 TB/s, with the measured HBM traffic
and issue counters of rocprofv3 --pmc child runs); the LBVH build time; the
p50 / p99 latency of a short serving run (64 resident meshes, Poisson
requests); and the reference CPU path (restated in oracle/port.cpp) timed on
this host at 1 thread and at this process's core share on a bounded sample of
the same rays.
"""
from __future__ import annotations

import argparse
import json
import os
import signal
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "raytracer-data-structures_amd"))
sys.path.insert(0, str(ROOT))

import rq  # noqa: E402
from rq import workloads as W  # noqa: E402

METRIC = "Mrays/s closest-hit (+% HBM roofline); LBVH build ms; serving p50/p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
MI355X_CUS, MI355X_XCDS = 256, 8
# LBVH build bytes per triangle of the reference byte model: Morton 56 (index +
# vertex gather in, code + index out), sort 68 (4 passes x 16 + histogram), Karras
# 28, refit 124
BUILD_MODEL_BYTES_PER_TRI = 276
# Algorithmic bytes of one LBVH build per triangle in the batch kernel (SCENE_AABB
# Morton map): bounds 48 (3 indices + 9 vertex floats), Morton 52 (same + code),
# sort 60 (pass 0: key in, key + index out; passes 1-3: key + index in and out),
# Karras 28 (key, two child links, two parent words, leaf range), refit 176
# (sorted index, 3 indices, 9 floats, 48 B triangle record, 64 B node boxes,
# arrival and height words).
BUILD_BYTES_PER_TRI = 48 + 52 + 60 + 28 + 176
# Traversal configuration per workload (closest variant, refill, any variant,
# refill, closest slice, any slice): the fastest of a sweep of every variant and
# refill threshold on MI355X (--sweep); fixed so that a profile of the same
# command times the same kernels. --autotune times the variants again instead.
TRACE_CONFIG = {"c1": None, "c2": "17,64,14,64,2048,1024", "c3": "17,16,17,16,2048,2048"}
# The LBVH rebuild of frame k+1 on a second stream (rq.FramePipeline): "frame"
# beside the whole of frame k, "tail" after frame k's last closest-hit launch,
# beside its shading only; None: sequential. Measured on MI355X (ms per frame,
# two runs each): --config c1 sequential 0.314, frame 0.219, tail 0.290;
# --config c2 sequential 3.097, frame 3.058, tail 3.051.
PIPELINE = {"c1": "frame", "c2": "tail"}
ALG_BASIS = ("algorithmic bytes of the canonical traversal (rays in, hits out, 32 B per node "
             "box tested, 36 B per triangle tested); node and triangle reads are mostly served "
             "from L2 and the Infinity Cache, so the fraction can exceed 1")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_share() -> int:
    """Host cores this process may run on, at most 16 (one GPU's share of the
    GPU box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def counted_batch(ctx, bvh, rays, any_hit: bool):
    """Traversal counts of the canonical traversal for one ray batch."""
    n = rays.n
    vis = rq.DeviceArray(ctx, 4 * n)
    tst = rq.DeviceArray(ctx, 4 * n)
    try:
        if any_hit:
            occ = rq.DeviceArray(ctx, n)
            ctx.any_hit_dev(bvh, rays, occ, counted=(vis, tst))
            ctx.sync()
            occ.free()
        else:
            hits = {k: rq.DeviceArray(ctx, 4 * n) for k in ("prim", "t")}
            ctx.closest_hit_dev(bvh, rays, hits, counted=(vis, tst))
            ctx.sync()
            for a in hits.values():
                a.free()
        v = vis.download(np.uint32, n).astype(np.int64)
        t = tst.download(np.uint32, n).astype(np.int64)
    finally:
        vis.free()
        tst.free()
    return int(v.sum()), int(t.sum())


# ----------------------------------------------------------- traversal config
def parse_trace_config(text: str) -> tuple[int, ...]:
    """closest variant, refill, any variant, refill and optionally the largest
    ray-stream slices (closest, any; default 1024)."""
    t = tuple(int(x) for x in text.split(","))
    if len(t) == 4:
        t += (1024, 1024)
    if len(t) != 6:
        raise SystemExit("--trace-config takes 4 or 6 comma-separated integers")
    return t


def apply_trace_config(ctx, cfg) -> None:
    vc, rc, va, ra, sc, sa = cfg
    ctx.trace_config(vc, va)
    ctx.trace_refill(rc, ra)
    ctx.trace_slice(sc, sa)


def choose_trace(ctx, bvh, closest_rays, any_rays, args, repeats: int):
    """--trace-config, else --autotune (time every variant on these rays), else
    the workload's fixed configuration (none: the context's defaults)."""
    if args.trace_config:
        cfg = parse_trace_config(args.trace_config)
    elif args.autotune:
        vc, rc, _ = ctx.autotune(bvh, closest_rays, any_hit=False, repeats=repeats)
        va, ra, _ = ctx.autotune(bvh, any_rays, any_hit=True, repeats=repeats)
        cfg = (vc, rc, va, ra) + ctx.trace_slice()
    elif TRACE_CONFIG.get(args.config):
        cfg = parse_trace_config(TRACE_CONFIG[args.config])
    else:
        return None
    apply_trace_config(ctx, cfg)
    return cfg


def trace_dict(ctx, cfg, how: str) -> dict | None:
    if not cfg:
        return None
    vc, rc, va, ra, sc, sa = cfg
    return {"chosen_by": how, "closest_variant": ctx.trace_variants()[vc],
            "any_variant": ctx.trace_variants()[va], "refill_idle": {"closest": rc, "any": ra},
            "slice": {"closest": sc, "any": sa}, "as_argument": ",".join(map(str, cfg))}


def cfg_arg(t) -> str | None:
    return ",".join(map(str, t)) if t else None


# ------------------------------------------------------- rocprofv3 counters
COPY_PROBE_BYTES = 1 << 30
KERNEL_SYMBOL = {"trace_closest": "k_trace<false, false", "trace_any": "k_trace<true, false",
                 "batch_build": "KernelMode)0", "batch_trace": "KernelMode)1"}
COPY_SYMBOL = "k_copy16"
# Three passes: HBM bytes read, HBM bytes written, and the issue / cache counters
# (at most 4 TCC, 8 SQ and 2 GRBM counters per pass).
PMC_PASSES = ("FETCH_SIZE", "WRITE_SIZE",
              "TCC_HIT_sum TCC_MISS_sum TCC_REQ_sum SQ_INSTS_VALU SQ_ACTIVE_INST_VALU "
              "SQ_THREAD_CYCLES_VALU SQ_WAVE_CYCLES SQ_WAIT_INST_ANY GRBM_GUI_ACTIVE")
CHILD_ARGS: list = []  # options every child run of this script repeats (set by main)


def copy_probe(ctx) -> None:
    """A library copy kernel over a known byte count: the calibration of the
    FETCH_SIZE / WRITE_SIZE / TCC_REQ counters."""
    src = rq.DeviceArray(ctx, COPY_PROBE_BYTES)
    dst = rq.DeviceArray(ctx, COPY_PROBE_BYTES)
    ctx.copy(dst, src, COPY_PROBE_BYTES)
    ctx.sync()
    src.free()
    dst.free()


def _pmc_run(counters: str, config: str, extra: list) -> dict | None:
    """One rocprofv3 --pmc child run of this script in probe mode (its own
    process group, removed whole afterwards): {counter: {kernel: (mean per
    dispatch, dispatches)}}."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(prof).exists():
        return None
    out = tempfile.mkdtemp(prefix="rq_pmc_", dir="/tmp")
    cmd = [prof, "--pmc", *counters.split(), "--output-format", "csv", "-d", out, "-o", "probe",
           "--", sys.executable, str(ROOT / "bench.py"), "--pmc-probe", "--config", config,
           "--no-cpu", "--no-serve", "--traffic", "off", "--warmup", "1", *CHILD_ARGS, *extra]
    env = dict(os.environ, TMPDIR="/tmp")
    proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                            stderr=subprocess.DEVNULL, start_new_session=True)
    try:
        rc = proc.wait(timeout=240)
    except subprocess.TimeoutExpired:
        rc = None
    finally:
        try:  # whatever the profiler left in its group
            os.killpg(proc.pid, signal.SIGKILL)
        except (ProcessLookupError, PermissionError):
            pass
        proc.wait()
    files = glob.glob(f"{out}/**/*counter_collection.csv", recursive=True)
    if not files:
        log(f"pmc {counters}: no counter file (rc {rc})")
        shutil.rmtree(out, ignore_errors=True)
        return None
    sums, counts = {}, {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            key = (row.get("Counter_Name"), row["Kernel_Name"])
            sums[key] = sums.get(key, 0.0) + float(row["Counter_Value"])
            counts[key] = counts.get(key, 0) + 1
    shutil.rmtree(out, ignore_errors=True)
    res = {}
    for (c, k), v in sums.items():
        res.setdefault(c, {})[k] = (v / counts[(c, k)], counts[(c, k)])
    return res


def _pick(per_kernel: dict, sym: str):
    """The mean of the most-dispatched kernel whose name contains sym."""
    hits = [(n, v) for k, (v, n) in per_kernel.items() if sym in k]
    return max(hits)[1] if hits else None


def measure_pmc(config: str, kernels, trace_cfg: str | None) -> dict:
    """Per kernel: HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, each scaled by
    known bytes / counter reading of the 1 GiB copy probe) and issue / cache
    figures: L2 hit rate; VALU issue (wave64 VALU instructions per CU per
    cycle, peak 1); lanes active per VALU instruction; the share of wave
    cycles spent waiting on memory counters; L2 request bytes per launch
    (TCC_REQ_sum scaled by the copy probe's bytes per request)."""
    extra = ["--trace-config", trace_cfg] if trace_cfg else []
    runs = [_pmc_run(p, config, extra) for p in PMC_PASSES]
    out = {}
    need = ("TCC_HIT_sum", "TCC_MISS_sum", "TCC_REQ_sum", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU",
            "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY", "GRBM_GUI_ACTIVE")
    for k in kernels:
        sym = KERNEL_SYMBOL[k]
        d = {}
        f, w, c = runs
        if f and w and "FETCH_SIZE" in f and "WRITE_SIZE" in w:
            kf, cf = _pick(f["FETCH_SIZE"], sym), _pick(f["FETCH_SIZE"], COPY_SYMBOL)
            kw, cw = _pick(w["WRITE_SIZE"], sym), _pick(w["WRITE_SIZE"], COPY_SYMBOL)
            if None not in (kf, cf, kw, cw) and cf and cw:
                sf, sw = (COPY_PROBE_BYTES / 1024.0) / cf, (COPY_PROBE_BYTES / 1024.0) / cw
                d["traffic"] = int(kf * 1024.0 * sf + kw * 1024.0 * sw)
                d["traffic_detail"] = {"fetch_kib": round(kf, 1), "fetch_scale": round(sf, 4),
                                       "write_kib": round(kw, 1), "write_scale": round(sw, 4)}
        if c and all(n in c for n in need):
            v = {n: _pick(c[n], sym) for n in need}
            if all(x is not None for x in v.values()):
                cycles = v["GRBM_GUI_ACTIVE"] / MI355X_XCDS
                copy_req = _pick(c["TCC_REQ_sum"], COPY_SYMBOL)
                req_bytes = (2.0 * COPY_PROBE_BYTES / copy_req) if copy_req else None
                d["counters"] = {
                    "l2_hit": round(v["TCC_HIT_sum"] / max(v["TCC_HIT_sum"] + v["TCC_MISS_sum"],
                                                           1.0), 4),
                    "valu_issue": round(v["SQ_INSTS_VALU"] / MI355X_CUS / max(cycles, 1.0), 4),
                    "valu_lane_util": round(v["SQ_THREAD_CYCLES_VALU"] /
                                            (64 * max(v["SQ_ACTIVE_INST_VALU"], 1.0)), 4),
                    "wave_wait_frac": round(v["SQ_WAIT_INST_ANY"] / max(v["SQ_WAVE_CYCLES"], 1.0),
                                            4),
                    "valu_insts": int(v["SQ_INSTS_VALU"]),
                    "l2_request_bytes": int(v["TCC_REQ_sum"] * req_bytes) if req_bytes else None,
                    "gpu_cycles": int(cycles),
                }
        out[k] = d
        log(f"pmc {k}: {d}")
    return out


def roofline_entry(kernel: str, alg: float, avg_ms: float, pmc: dict | None) -> dict:
    """The roofline object of one kernel: algorithmic bytes per launch / mean
    launch time against the HBM peak, the measured HBM traffic of the same
    launch and what it is as a fraction of the peak, and the issue counters
    (the binding limit is the larger of valu_issue and hbm_traffic_frac)."""
    achieved = alg / (avg_ms * 1e-3) / 1e9
    r = dict(kernel=kernel, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
             unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
             basis=ALG_BASIS, bytes_per_launch=int(alg), avg_launch_ms=round(avg_ms, 4))
    pmc = pmc or {}
    if pmc.get("traffic") is not None:
        r["traffic"] = pmc["traffic"]
        r["traffic_detail"] = pmc["traffic_detail"]
        r["hbm_traffic_frac"] = round(pmc["traffic"] / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if pmc.get("counters"):
        c = pmc["counters"]
        r["counters"] = c
        lim = {"valu_issue": c["valu_issue"], "hbm": r.get("hbm_traffic_frac", 0.0)}
        r["limiter"] = max(lim, key=lim.get)
    return r


# ------------------------------------------------------------- CPU baseline
CPU_SAMPLE_RAYS = 1_000_000  # rays of one timed pass (about 1.5 s at 1 thread)


def _median_time(fn, runs: int, warmup: bool = True) -> tuple[float, list]:
    if warmup:
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 4) for t in ts]


def cpu_baseline(mesh, batches, what: str, target_rays: int = CPU_SAMPLE_RAYS,
                 build_runs: int = 3, build_warmup: bool = True, query_runs: int = 3) -> dict:
    """The reference CPU path restated on triangles (oracle/port.cpp), timed on
    this host with steady wall clocks: the reference's median-split BVH build
    (constructBVHNew), its "LBVH" build (constructLBVHTree), and this build's
    CPU LBVH (Morton, stable sort, Karras, refit) next to the GPU build; then
    collect-then-test closest hit on every k-th active ray of each batch (k
    chosen for about target_rays rays) at 1 thread, as the reference's render
    loop, and on this process's core share. Each: 1 warm-up, median of 3."""
    from oracle import oracle as O

    cores = cpu_share()
    builds = {}
    for name, fn in (("reference_bvh_constructBVHNew", lambda: O.PortBVH(mesh)),
                     ("reference_lbvh_constructLBVHTree", lambda: O.PortBVH(mesh, lbvh=True)),
                     ("oracle_lbvh", lambda: O.OracleBVH(mesh, rq.MORTON_SCENE_AABB))):
        med, ts = _median_time(fn, build_runs, build_warmup)
        builds[name] = {"ms": round(med * 1e3, 1), "runs_s": ts}
    active = sum(int((h["tmin"] < h["tmax"]).sum()) for h in batches)
    every = max(1, active // target_rays)
    sample = {k: [] for k in O.RAY_FIELDS}
    for host in batches:
        idx = np.nonzero(host["tmin"] < host["tmax"])[0][::every]
        for k in O.RAY_FIELDS:
            sample[k].append(host[k][idx])
    sample = {k: np.concatenate(v) for k, v in sample.items()}
    n = len(sample["ox"])
    port = O.PortBVH(mesh)
    t1, ts1 = _median_time(lambda: port.closest(sample, threads=1), query_runs)
    tn, tsn = _median_time(lambda: port.closest(sample, threads=cores), query_runs)
    timing = f"1 warm-up, median of {query_runs} (builds: " + (
        f"1 warm-up, median of {build_runs})" if build_warmup else f"{build_runs} run)")
    return dict(value=round(n / tn / 1e6, 4), unit="Mrays/s", cores=cores, kind="port",
                sample=f"every {every}th active ray {what}: {n} rays per pass, closest hit",
                rays=n, single_thread={"value": round(n / t1 / 1e6, 4), "cores": 1,
                                       "runs_s": ts1},
                all_cores_runs_s=tsn, builds=builds, timing=timing,
                host=f"{cpu_model()}, {os.cpu_count()} logical CPUs visible, {cores} used")


# ------------------------------------------------------------------ serving
def serve_run(ctx, warm: float, meas: float, saturation: bool) -> dict:
    """Serving: 64 resident meshes (F ~ log-U[1e4, 2e6], seed 5) in a server
    with 72 slots of 2^21 triangles (64 resident, 8 for uploads) and a 200 GB
    HBM budget. Open loop: Poisson arrivals at 2000 requests/s for `warm`
    seconds of warm-up plus `meas` measured; request sizes log-U[4096,
    4194304] rays (seed 7), alternately pinhole rays aimed at the mesh and
    uniform random rays (generated on the device from the request's
    parameters), closest hit with prim + t returned to page-locked host
    arrays; 5% of requests upload a new mesh (from a pool of 8 pre-generated
    meshes, seed 6) that is built before its rays are traced. Latency =
    results on the host - scheduled arrival. With `saturation`, a burst of the
    first 1000 non-upload requests measures the served throughput, and the
    same requests go through the plain per-request path (ray generation,
    rq_closest_hit, copy back, wait)."""
    import ctypes as C

    t_setup = time.perf_counter()
    resident = W.c5_meshes(64, 5)
    uploads = W.c5_meshes(8, 6)
    bounds = {id(m): m.bounds() for m in resident + uploads}
    log(f"serve: {sum(m.ntri for m in resident)} resident triangles, "
        f"{sum(m.ntri for m in uploads)} in the upload pool ({time.perf_counter() - t_setup:.1f} s)")
    srv = rq.Server(0, mesh_slots=len(resident) + len(uploads))
    slots = [srv.load_mesh(m) for m in resident]
    slot_mesh = dict(zip(slots, resident))
    lam = 2000.0
    rng = np.random.default_rng(7)
    n_req = int(lam * (warm + meas) * 1.2) + 16
    arrival = np.cumsum(rng.exponential(1.0 / lam, n_req))
    n_req = int(np.searchsorted(arrival, warm + meas))
    arrival = arrival[:n_req]
    sizes = np.rint(np.exp(rng.uniform(np.log(4096.0), np.log(4194304.0), n_req))).astype(np.int64)
    target = rng.integers(0, len(slots), n_req)
    upload = rng.random(n_req) < 0.05
    max_n = int(sizes.max())
    ring = 80  # page-locked result buffers (at most 80 requests in flight from this load generator)
    bufs = [(rq.HostArray(max_n, np.int32), rq.HostArray(max_n, np.float32)) for _ in range(ring)]
    log(f"serve: {n_req} requests over {warm + meas:.1f} s, {int(upload.sum())} uploads, "
        f"offered {sizes.sum() / (warm + meas) / 1e6:.0f} Mrays/s")

    def spec(j):
        if upload[j]:
            m = uploads[j % len(uploads)]
            sp = W.c5_rays_spec(bounds[id(m)], j % 2, 100000 + j)
            sp["new_mesh"] = m
        else:
            m = slot_mesh[slots[target[j]]]
            sp = W.c5_rays_spec(bounds[id(m)], j % 2, 100000 + j)
            sp["mesh"] = slots[target[j]]
        return sp

    def drive(order, t0, times):
        """Submits order[k] at t0 + times[k] (or as soon as a buffer is free);
        returns per-request (t_done - scheduled arrival, steps)."""
        free = list(range(ring))
        inflight = {}
        lat = np.full(len(order), np.nan)
        steps = np.zeros(len(order), np.int64)
        k = 0
        while k < len(order) or inflight:
            now = rq.server_clock()
            while k < len(order) and free and t0 + times[k] <= now:
                j = order[k]
                b = free.pop()
                tk = srv.submit(int(sizes[j]), rq.Q_CLOSEST, prim=bufs[b][0].array,
                                t=bufs[b][1].array, **spec(j))
                inflight[tk] = (k, b)
                k += 1
            for tk, st in srv.take_done():
                kk, b = inflight.pop(tk)
                if st.state != rq.REQ_DONE:
                    raise RuntimeError(f"request {order[kk]} failed: {st.state}")
                lat[kk] = st.t_done - (t0 + times[kk])
                steps[kk] = st.steps
                free.append(b)
                srv.forget(tk)
        return lat, steps

    s0 = srv.stats()
    t0 = rq.server_clock() + 0.05
    lat, steps = drive(list(range(n_req)), t0, arrival)
    wall = rq.server_clock() - t0
    s1 = srv.stats()
    sel = arrival >= warm
    lat_ms = lat[sel] * 1e3
    up_ms = lat[sel & upload] * 1e3
    p50, p99 = (float(np.percentile(lat_ms, q)) for q in (50, 99))
    log(f"serve: p50 {p50:.2f} ms, p99 {p99:.2f} ms, wall {wall:.2f} s")
    out = {
        "resident_meshes": len(slots), "resident_triangles": int(sum(m.ntri for m in resident)),
        "arrival_rate_per_s": lam, "warmup_s": warm, "measured_s": meas,
        "requests": int(sel.sum()), "uploads": int((sel & upload).sum()),
        "hbm_budget_bytes": 200 * 10 ** 9,
        "offered_mrays_s": round(float(sizes[sel].sum()) / meas / 1e6, 1),
        "query": "closest (prim, t) to host", "rays": "half pinhole, half uniform",
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3),
                       "mean": round(float(np.mean(lat_ms)), 3),
                       "max": round(float(np.max(lat_ms)), 3),
                       "upload_p50": round(float(np.percentile(up_ms, 50)), 3) if len(up_ms) else None,
                       "upload_p99": round(float(np.percentile(up_ms, 99)), 3) if len(up_ms) else None},
        "trace_launches": int(s1["steps"] - s0["steps"]),
        "mean_launches_per_request": round(float(steps[sel].mean()), 2),
        "max_requests_per_launch": int(s1["max_step_requests"]),
        "gpu_trace_busy_fraction": round((s1["busy_s"] - s0["busy_s"]) / wall, 3),
    }
    if saturation:
        burst = [j for j in range(n_req) if not upload[j]][:1000]
        t0 = rq.server_clock() + 0.01
        blat, _ = drive(burst, t0, np.zeros(len(burst)))
        served_s = float(np.nanmax(blat))
        burst_rays = int(sizes[burst].sum())
        out["server_stats"] = {k: (round(v, 4) if isinstance(v, float) else v)
                               for k, v in srv.stats().items()}
        srv.close()
        lib = rq.lib()
        bvhs = {}
        for sl, m in slot_mesh.items():
            sc = ctx.upload_scene(m)
            bvhs[sl] = ctx.build_lbvh(sc, rq.MORTON_SCENE_AABB)
        dr = rq.DeviceRays(ctx, max_n)
        dprim, dt = rq.DeviceArray(ctx, 4 * max_n), rq.DeviceArray(ctx, 4 * max_n)
        ptrs = [dr.arrays[f].ptr for f in rq.RAY_FIELDS]
        rays_st = dr.struct()

        def plain(j):
            n = int(sizes[j])
            sp = spec(j)
            if "pinhole" in sp:
                p = sp["pinhole"]
                cam = rq.Pinhole((C.c_float * 3)(*p["eye"]), (C.c_float * 3)(*p["look_at"]),
                                 (C.c_float * 3)(*p["up"]), p["fov_deg"], p["width"])
                ctx._check(lib.rq_raygen_pinhole(ctx.h, C.byref(cam), n, sp["seed"], *ptrs, None))
            else:
                lo3 = (C.c_float * 3)(*sp["box"][0])
                hi3 = (C.c_float * 3)(*sp["box"][1])
                ctx._check(lib.rq_raygen_uniform(ctx.h, n, sp["seed"], lo3, hi3, *ptrs, None))
            h = rq.Hits(dprim.ptr, dt.ptr, None, None)
            ctx._check(lib.rq_closest_hit(ctx.h, bvhs[sp["mesh"]], C.byref(rays_st), n,
                                          C.byref(h), None))
            ctx._check(lib.rq_memcpy_d2h(ctx.h, bufs[0][0].ptr, dprim.ptr, 4 * n, None))
            ctx._check(lib.rq_memcpy_d2h(ctx.h, bufs[0][1].ptr, dt.ptr, 4 * n, None))

        for j in burst[:20]:
            plain(j)
        ctx.sync()
        t_a = time.perf_counter()
        for j in burst:
            plain(j)
        ctx.sync()
        plain_s = time.perf_counter() - t_a
        served_mrays, plain_mrays = burst_rays / served_s / 1e6, burst_rays / plain_s / 1e6
        out["saturation"] = {"requests": len(burst), "rays": burst_rays,
                             "served_mrays_s": round(served_mrays, 1),
                             "plain_per_request_mrays_s": round(plain_mrays, 1),
                             "speedup": round(served_mrays / plain_mrays, 3)}
        dr.free()
        dprim.free()
        dt.free()
        for b in bvhs.values():
            ctx.release(b)
    else:
        srv.close()
    for a, b in bufs:
        a.free()
        b.free()
    return out


# -------------------------------------------------------------- workloads
def base_line(value, ms_per_step, steps, warmup, config) -> dict:
    return {"metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": config}


def run_frames(ctx, args) -> int:
    """--config c1 / c2: rebuild + one frame of the wavefront renderer per step."""
    mesh, params, workload = (W.c2 if args.config == "c2" else W.c1)()
    log(f"{workload}: {mesh.ntri} triangles")
    scene = ctx.upload_scene(mesh)
    bvh = ctx.build_lbvh(scene, rq.MORTON_SCENE_AABB)
    if args.treelets:
        ctx.treelets(bvh, args.treelets)
    ren = rq.Renderer(ctx, params)
    cfg = None
    if not args.autotune:  # fixed configuration: every launch, warm-up included, uses it
        cfg = choose_trace(ctx, None, None, None, args, 0)

    # Every frame rebuilds its LBVH; pipelined, the rebuild for frame k+1 runs
    # on its own stream while frame k renders (rq.FramePipeline)
    pipe = rq.FramePipeline(ctx, ren, scene, bvh, rq.MORTON_SCENE_AABB, args.treelets,
                            pipelined=args.pipeline is not None,
                            overlap=args.pipeline or "frame")
    frame = pipe.frame

    for _ in range(args.warmup):
        frame()
    ctx.sync()
    cnt = ren.counts()
    rays_per_frame = int(cnt.primary + cnt.bounce + cnt.shadow)
    info = ctx.bvh_info(bvh)
    batches = {"primary": ren.rays(0), "shadow": ren.rays(2)}
    if params.bounce:
        batches["bounce"] = ren.rays(1)
    counts = {k: counted_batch(ctx, bvh, r, any_hit=(k == "shadow")) for k, r in batches.items()}
    if args.sweep:
        sweep(ctx, bvh, batches)
    how = "--trace-config" if args.trace_config else "fixed for this workload"
    if args.autotune:
        cfg = choose_trace(ctx, bvh, batches["primary"], batches["shadow"], args, 3)
        how = "autotuned on this frame's primary and shadow rays"
        for _ in range(2):
            frame()
        ctx.sync()
    if args.pmc_probe:
        for _ in range(2):
            frame()
        ctx.sync()
        copy_probe(ctx)
        return 0

    # ---- timed region: K frames, HIP events on the launch streams around the
    # traversal launches only (events around every kernel cost ~4 % of a 1.31 M-triangle frame)
    ctx.profile_only(["trace_closest", "trace_any"])
    ctx.profile(True)
    ctx.profile_reset()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.profile(False)
    prof = ctx.profile_read()
    assert ctx.device_errors() == 0
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = rays_per_frame * args.steps / elapsed / 1e6

    # ---- per-kernel breakdown (build and shading included) from frames after
    # the timed region, every kernel bracketed by events
    breakdown_frames = 3
    ctx.profile_only(None)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(breakdown_frames):
        frame()
    ctx.sync()
    ctx.profile(False)
    breakdown = ctx.profile_read()

    # ---- rooflines: algorithmic bytes per launch / event time
    n = ren.n
    closest_rays = [("primary", n)] + ([("bounce", n)] if params.bounce else [])
    closest_bytes = sum(r * (32 + 8) + 32 * counts[k][0] + 36 * counts[k][1]
                        for k, r in closest_rays) / len(closest_rays)
    nshadow = batches["shadow"].n
    any_bytes = nshadow * (32 + 1) + 32 * counts["shadow"][0] + 36 * counts["shadow"][1]
    # the shadow rays may go out in several launches per frame (primary hits'
    # and second hits' concurrently): bytes per launch = frame bytes / launches
    any_launches = max(1, prof.get("trace_any", (0, args.steps))[1] // args.steps)
    alg = {"trace_closest": closest_bytes, "trace_any": any_bytes / any_launches}
    avg = {k: prof[k][0] / prof[k][1] for k in alg if k in prof}
    dom = max(avg, key=lambda k: prof[k][0])
    pmc = {}
    if args.traffic == "auto":
        pmc = measure_pmc(args.config, list(avg), cfg_arg(cfg))
    roofline = roofline_entry(dom, alg[dom], avg[dom], pmc.get(dom))
    roof_closest = (roofline_entry("trace_closest", alg["trace_closest"], avg["trace_closest"],
                                   pmc.get("trace_closest")) if "trace_closest" in avg else None)
    roof_any = (roofline_entry("trace_any", alg["trace_any"], avg["trace_any"], pmc.get("trace_any"))
                if "trace_any" in avg else None)
    build_kernels = ("build_init", "centroid_bounds", "morton_hist", "radix_onesweep", "karras",
                     "refit", "refit_top", "widen")
    build_in_frame_ms = (sum(breakdown[k][0] for k in build_kernels if k in breakdown) /
                         breakdown_frames)
    # the build alone (pipelined, the per-kernel times above include the
    # frame's work beside it): three rebuilds one after the other
    build_ms = build_in_frame_ms
    if args.pipeline:
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(3):
            ctx.rebuild(bvh)
        ctx.sync()
        ctx.profile(False)
        alone = ctx.profile_read()
        build_ms = sum(alone[k][0] for k in build_kernels if k in alone) / 3
    closest_ms = prof.get("trace_closest", (0, 1))[0] / args.steps
    closest_rays_n = sum(r for _, r in closest_rays)

    line = base_line(value, ms_per_step, args.steps, args.warmup, {
        "workload": workload, "triangles": mesh.ntri,
        "width": params.width, "height": params.height, "spp": params.spp,
        "bounce": int(params.bounce), "rays_per_frame": rays_per_frame,
        "rays_primary": int(cnt.primary), "rays_bounce": int(cnt.bounce),
        "rays_shadow": int(cnt.shadow), "lbvh_rebuilt_per_frame": True,
        "rebuild_overlaps_previous_frame": args.pipeline or False,
        "bvh_height": int(info.height)})
    line["lbvh_build_ms"] = round(build_ms, 4)
    line["lbvh_build_ms_beside_frame"] = round(build_in_frame_ms, 4) if args.pipeline else None
    line["lbvh_build_roofline"] = {
        "bytes_per_tri": BUILD_MODEL_BYTES_PER_TRI,
        "achieved_gbs": round(mesh.ntri * BUILD_MODEL_BYTES_PER_TRI / (build_ms * 1e-3) / 1e9, 1),
        "frac": round(mesh.ntri * BUILD_MODEL_BYTES_PER_TRI / (build_ms * 1e-3) / 1e9 /
                      HBM_PEAK_GBS, 4)}
    line["closest_hit_mrays_s"] = (round(closest_rays_n / (closest_ms * 1e-3) / 1e6, 1)
                                   if closest_ms else None)
    line["kernel_ms_per_step"] = {k: round(v[0] / breakdown_frames, 4)
                                  for k, v in breakdown.items()}
    line["traversal_counts"] = {k: {"node_visits": v[0], "tri_tests": v[1]}
                                for k, v in counts.items()}
    line["trace_config"] = trace_dict(ctx, cfg, how)
    line["roofline"] = roofline
    line["roofline_closest"] = roof_closest
    line["roofline_any"] = roof_any
    line["serving"] = None
    line["cpu_baseline"] = None
    if args.config == "c2" and not args.no_serve:
        line["serving"] = serve_run(ctx, 0.5, args.serve_seconds_short, saturation=False)
    if not args.no_cpu:
        host = [batches[k].download() for k in batches]
        log(f"cpu baseline: 1 and {cpu_share()} threads")
        line["cpu_baseline"] = cpu_baseline(
            mesh, host, "of each batch of one frame (primary, bounce, shadow as closest hit)",
            target_rays=args.cpu_rays)
    print(json.dumps(line), flush=True)
    ren.close()
    ctx.close()
    return 0


def run_c3(ctx, args) -> int:
    """Raw ray batch: 256 displaced icospheres of 20,480 triangles (5,242,880)
    with Poisson-disk centres in [0,100]^3; 2^24 rays with origins uniform in
    the scene box inflated by 10% per side and directions uniform on the
    sphere; one step = closest hit + any hit over the same rays."""
    mesh, (lo, hi), n, ray_seed = W.c3()
    log(f"c3_raw_ray_batch: {mesh.ntri} triangles")
    scene = ctx.upload_scene(mesh)
    bvh = ctx.build_lbvh(scene, rq.MORTON_SCENE_AABB)
    treelets = None
    if args.treelets:  # the BVH is built once: time the restructuring once, on its own
        ctx.sync()
        ctx.profile(True)
        ctx.profile_reset()
        ctx.treelets(bvh, args.treelets)
        ctx.sync()
        ctx.profile(False)
        tp = ctx.profile_read()
        treelets = {"rounds": args.treelets,
                    "ms": round(sum(v[0] for k, v in tp.items() if k.startswith("treelet")), 3),
                    "height": ctx.bvh_info(bvh).height}
        log(f"treelets: {treelets}")
    rays = ctx.raygen_uniform(n, ray_seed, lo, hi)
    hits = {k: rq.DeviceArray(ctx, 4 * n) for k in ("prim", "t")}
    occ = rq.DeviceArray(ctx, n)
    vc, tc = counted_batch(ctx, bvh, rays, any_hit=False)
    va, ta = counted_batch(ctx, bvh, rays, any_hit=True)
    if args.sweep:
        sweep(ctx, bvh, {"closest": rays, "any": rays})
    cfg = choose_trace(ctx, bvh, rays, rays, args, 2)
    how = ("--trace-config" if args.trace_config else
           "autotuned on these rays" if args.autotune else "fixed for this workload")
    for _ in range(args.warmup):
        ctx.closest_hit_dev(bvh, rays, hits)
        ctx.any_hit_dev(bvh, rays, occ)
    ctx.sync()
    if args.pmc_probe:
        ctx.closest_hit_dev(bvh, rays, hits)
        ctx.any_hit_dev(bvh, rays, occ)
        ctx.sync()
        copy_probe(ctx)
        return 0
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.closest_hit_dev(bvh, rays, hits)
        ctx.any_hit_dev(bvh, rays, occ)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.profile(False)
    prof = ctx.profile_read()
    assert ctx.device_errors() == 0
    elapsed = t1 - t0
    prim = hits["prim"].download(np.int32, n)
    occ_h = occ.download(np.uint8, n)
    avg = {k: prof[k][0] / prof[k][1] for k in ("trace_closest", "trace_any")}
    alg = {"trace_closest": n * (32 + 8) + 32 * vc + 36 * tc,
           "trace_any": n * (32 + 1) + 32 * va + 36 * ta}
    dom = max(avg, key=avg.get)
    pmc = measure_pmc("c3", list(avg), cfg_arg(cfg)) if args.traffic == "auto" else {}
    line = base_line(2 * n * args.steps / elapsed / 1e6, elapsed / args.steps * 1e3, args.steps,
                     args.warmup, {"workload": "c3_raw_ray_batch", "triangles": mesh.ntri,
                                   "rays": n, "queries": "closest+any",
                                   "closest_hit_rate": round(float((prim >= 0).mean()), 4),
                                   "occluded_rate": round(float(occ_h.mean()), 4)})
    line["closest_hit_mrays_s"] = round(n / (avg["trace_closest"] * 1e-3) / 1e6, 1)
    line["any_hit_mrays_s"] = round(n / (avg["trace_any"] * 1e-3) / 1e6, 1)
    line["kernel_ms_per_step"] = {k: round(v[0] / args.steps, 4) for k, v in prof.items()}
    line["traversal_counts"] = {"closest": {"node_visits": vc, "tri_tests": tc},
                                "any": {"node_visits": va, "tri_tests": ta}}
    line["trace_config"] = trace_dict(ctx, cfg, how)
    line["roofline"] = roofline_entry(dom, alg[dom], avg[dom], pmc.get(dom))
    line["roofline_closest"] = roofline_entry("trace_closest", alg["trace_closest"],
                                              avg["trace_closest"], pmc.get("trace_closest"))
    line["cpu_baseline"] = None
    if treelets:
        line["lbvh_treelets"] = treelets
    if not args.no_cpu:
        host = rays.download()
        line["cpu_baseline"] = cpu_baseline(mesh, [host], "of the 2^24 uniform rays",
                                            target_rays=args.cpu_rays, build_runs=1,
                                            build_warmup=False)
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def c4_workload(ctx, count: int = 512):
    """The batch's meshes (uploaded) and device rays: uniform in each mesh's
    box inflated by 10% per side, directions uniform on the sphere."""
    tris, nrays = W.c4_sizes(count)
    meshes, scenes, rays = [], [], []
    for m in range(count):
        mesh = W.c4_mesh(m, tris[m])
        lo, hi = W.inflated_box(mesh)
        scenes.append(ctx.upload_scene(mesh))
        rays.append(ctx.raygen_uniform(int(nrays[m]), W.c4_ray_seed(m), lo, hi))
        meshes.append(mesh)
    return meshes, scenes, rays


def run_c4(ctx, args) -> int:
    """512 meshes built and traced (closest + any hit) in one persistent launch
    replayed from a hipGraph, against the plain per-request path (per mesh:
    LBVH rebuild, closest-hit launch, any-hit launch, wait) on the same
    device-resident meshes, rays and output buffers."""
    t_setup = time.perf_counter()
    meshes, scenes, rays = c4_workload(ctx)
    ntri = sum(m.ntri for m in meshes)
    nray = sum(r.n for r in rays)
    log(f"c4: {len(meshes)} meshes, {ntri} triangles, {nray} rays "
        f"(setup {time.perf_counter() - t_setup:.1f} s)")
    hits = [{k: rq.DeviceArray(ctx, 4 * r.n) for k in ("prim", "t")} for r in rays]
    occ = [rq.DeviceArray(ctx, r.n) for r in rays]
    items = [dict(scene=s, rays=r, hits=h, occ=o) for s, r, h, o in zip(scenes, rays, hits, occ)]
    batch = rq.Batch(ctx, items, rq.MORTON_SCENE_AABB)
    tuned = None
    if args.autotune:
        v, tile, refill, ms = batch.autotune(1)
        tuned = {"variant": rq.Batch.configs()[v], "ray_tile": tile, "refill_idle": refill,
                 "ms": round(ms, 3)}
        tuned["build_waves"] = batch.build_waves()
        log(f"batch autotune: {tuned}")
    batch_cfg = batch.config()
    for _ in range(max(args.warmup, 1)):
        batch.run_graph()
    assert batch.errors() == 0, "batch errors"
    if args.pmc_probe:
        batch.run_graph()
        copy_probe(ctx)
        return 0

    # ---- batched: K replays of the captured persistent launch
    ctx.profile(True)
    ctx.profile_reset()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run_graph()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.profile(False)
    prof = ctx.profile_read()
    assert batch.errors() == 0, "batch errors"
    batch_s = (t1 - t0) / args.steps
    # the run's two kernels (build phases + widen, then the trace), event-timed
    # on plain runs after the timed replays
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(3):
        batch.run()
    ctx.sync()
    ctx.profile(False)
    split = ctx.profile_read()
    phase_ms = {k: round(split[f"batch_{k}"][0] / split[f"batch_{k}"][1], 3)
                for k in ("build", "trace") if f"batch_{k}" in split}
    log(f"batch kernels: {phase_ms}")
    check = [0, 1, len(meshes) // 2, len(meshes) - 1]
    ref = {m: (hits[m]["prim"].download(np.int32, rays[m].n), occ[m].download(np.uint8, rays[m].n))
           for m in check}

    # ---- plain per-request path: what a caller without the batch API does,
    # per mesh a standalone LBVH (allocated here, rebuilt in the timed loop)
    bvhs = [ctx.build_lbvh(scenes[m], rq.MORTON_SCENE_AABB) for m in range(len(meshes))]
    ctx.sync()
    if args.autotune:
        med = int(np.argsort([r.n for r in rays])[len(rays) // 2])
        choose_trace(ctx, bvhs[med], rays[med], rays[med], args, 2)

    def per_request():
        for m in range(len(meshes)):
            ctx.rebuild(bvhs[m])
            ctx.closest_hit_dev(bvhs[m], rays[m], hits[m])
            ctx.any_hit_dev(bvhs[m], rays[m], occ[m])
            ctx.sync()

    per_request()
    base_steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(base_steps):
        per_request()
    t1 = time.perf_counter()
    plain_s = (t1 - t0) / base_steps
    # its kernel time alone (rebuild kernels + the two traversals per mesh),
    # event-timed on one more pass
    ctx.profile(True)
    ctx.profile_reset()
    per_request()
    ctx.profile(False)
    pr = ctx.profile_read()
    per_request_kernels = {k: round(v[0], 3) for k, v in pr.items()}
    log(f"per-request kernels (ms per pass): {per_request_kernels}")
    assert ctx.device_errors() == 0
    for m in check:
        assert np.array_equal(hits[m]["prim"].download(np.int32, rays[m].n), ref[m][0]), m
        assert np.array_equal(occ[m].download(np.uint8, rays[m].n), ref[m][1]), m

    # ---- roofline of the batch kernel: build bytes + traversal bytes
    vis_c = tst_c = vis_a = tst_a = 0
    for m in range(len(meshes)):
        a, b = counted_batch(ctx, bvhs[m], rays[m], any_hit=False)
        c, d = counted_batch(ctx, bvhs[m], rays[m], any_hit=True)
        vis_c, tst_c, vis_a, tst_a = vis_c + a, tst_c + b, vis_a + c, tst_a + d
    alg_trace = (nray * (32 + 8) + 32 * vis_c + 36 * tst_c + nray * (32 + 1) + 32 * vis_a +
                 36 * tst_a)
    alg_build = BUILD_BYTES_PER_TRI * ntri
    pmc = measure_pmc("c4", ["batch_trace", "batch_build"], None) if args.traffic == "auto" else {}
    line = base_line(2 * nray / batch_s / 1e6, batch_s * 1e3, args.steps, args.warmup, {
        "workload": "c4_batched_build_and_trace", "meshes": len(meshes), "triangles": ntri,
        "rays": nray, "queries": "build+closest+any",
        "launch": "build kernel (every mesh's build phases + widen) then trace kernel "
                  "(4-wide walk), one hipGraph replay per step",
        "batch_config": {"trace": rq.Batch.configs()[batch_cfg[0]], "ray_tile": batch_cfg[1],
                         "refill_idle": batch_cfg[2], "build_waves": batch.build_waves()}})
    line["batch_kernel_ms"] = phase_ms
    line["per_request_kernel_ms"] = round(sum(per_request_kernels.values()), 3)
    line["per_request_trace_ms"] = round(sum(v for k, v in per_request_kernels.items()
                                             if k.startswith("trace_")), 3)
    line["per_request_ms"] = round(plain_s * 1e3, 3)
    line["per_request_mrays_s"] = round(2 * nray / plain_s / 1e6, 2)
    line["speedup_vs_per_request"] = round(plain_s / batch_s, 3)
    line["traversal_counts"] = {"closest": {"node_visits": vis_c, "tri_tests": tst_c},
                                "any": {"node_visits": vis_a, "tri_tests": tst_a}}
    line["autotune"] = tuned
    # the dominant kernel: the trace kernel (closest + any hit of every mesh)
    line["roofline"] = roofline_entry("k_batch trace (4-wide walk)", alg_trace,
                                      phase_ms["trace"], pmc.get("batch_trace"))
    line["roofline"]["basis"] = ALG_BASIS
    line["build_roofline"] = roofline_entry("k_batch build phases", alg_build, phase_ms["build"],
                                            pmc.get("batch_build"))
    line["build_roofline"]["basis"] = (f"build bytes ({BUILD_BYTES_PER_TRI} per triangle: "
                                       "the build kernels' reads and writes)")
    line["cpu_baseline"] = None
    if not args.no_cpu:
        sample = list(range(0, len(meshes), 61))  # icospheres and soups
        parts = [cpu_baseline(meshes[m], [rays[m].download()], "", target_rays=200_000,
                              build_runs=1, build_warmup=False, query_runs=1) for m in sample]
        rays_n = sum(p["rays"] for p in parts)
        t_n = sum(p["rays"] / p["value"] for p in parts)
        t_1 = sum(p["rays"] / p["single_thread"]["value"] for p in parts)
        line["cpu_baseline"] = dict(
            value=round(rays_n / t_n, 4), unit="Mrays/s", cores=parts[0]["cores"], kind="port",
            sample=f"meshes {sample}: about 200,000 rays of each (closest hit), 1 warm-up and "
                   f"1 timed pass; builds once each",
            single_thread={"value": round(rays_n / t_1, 4), "cores": 1},
            builds_ms={k: round(sum(p["builds"][k]["ms"] for p in parts), 1)
                       for k in parts[0]["builds"]},
            host=parts[0]["host"])
    print(json.dumps(line), flush=True)
    batch.close()
    ctx.close()
    return 0


def run_c5(ctx, args) -> int:
    """Serving (serve_run) for --serve-seconds, with the saturation burst and
    the plain per-request comparison."""
    s = serve_run(ctx, 1.0, float(args.serve_seconds), saturation=True)
    keys = ("resident_meshes", "resident_triangles", "arrival_rate_per_s", "measured_s",
            "requests", "uploads", "hbm_budget_bytes", "offered_mrays_s", "query", "rays")
    line = base_line(s["offered_mrays_s"], s["latency_ms"]["mean"], s["requests"], 0,
                     {"workload": "c5_serving", **{k: s[k] for k in keys}})
    line.update({k: v for k, v in s.items() if k not in keys})
    line["roofline"] = None
    line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def sweep(ctx, bvh, batches) -> None:
    """Time each traversal variant / refill threshold on this frame's batches."""
    saved = (ctx.trace_config()[:2], ctx.trace_refill())
    for v in ctx.trace_variants():
        for refill in (16, 48, 56, 64):
            row = []
            for name, rays in batches.items():
                anyh = name in ("shadow", "any")
                if anyh and v["tstack"]:
                    row.append(f"{name}=   -   ")
                    continue
                ctx.trace_config(v["index"], v["index"], refill)
                out = rq.DeviceArray(ctx, 4 * rays.n)
                tt = rq.DeviceArray(ctx, 4 * rays.n)
                ctx.profile(True)
                ctx.profile_reset()
                for _ in range(4):
                    if anyh:
                        ctx.any_hit_dev(bvh, rays, out)
                    else:
                        ctx.closest_hit_dev(bvh, rays, {"prim": out, "t": tt})
                ctx.sync()
                prof = ctx.profile_read()
                ctx.profile(False)
                key = "trace_any" if anyh else "trace_closest"
                ms, cnt = prof[key]
                row.append(f"{name}={ms / cnt:7.3f}")
                out.free()
                tt.free()
            log(f"variant {v['index']} b{v['block']} d{v['lds_depth']} t{v['tstack']} "
                f"refill {refill:2d}: " + " ".join(row))
    ctx.trace_config(*saved[0])
    ctx.trace_refill(*saved[1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 steps: a pipelined frame's last rebuild (not overlapped by a next
    # frame) adds build / steps to the step time
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="c5: measured seconds of arrivals (after 1 s of warm-up)")
    ap.add_argument("--serve-seconds-short", type=float, default=2.0,
                    help="c2: measured seconds of the serving run reported beside the frame")
    ap.add_argument("--no-serve", action="store_true", help="c2: no serving run")
    ap.add_argument("--pipeline", default="auto", choices=["auto", "frame", "tail", "off"],
                    help="c1/c2: rebuild the next frame's LBVH while a frame renders: beside "
                         "the whole frame (frame), beside its shading after its last "
                         "closest-hit launch (tail), or not (off); auto: the workload's choice")
    ap.add_argument("--traffic", default="auto", choices=["auto", "off"],
                    help="HBM bytes and issue counters per launch from rocprofv3 --pmc child runs")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="(child of --traffic) run the workload briefly, then a 1 GiB copy")
    ap.add_argument("--cpu-rays", type=int, default=CPU_SAMPLE_RAYS,
                    help="rays of one timed CPU-baseline pass")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", action="store_true",
                    help="time every traversal variant on this frame's batches (stderr)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the traversal variants on the workload's rays instead of "
                         "the fixed configuration")
    ap.add_argument("--treelets", type=int, default=0, metavar="R",
                    help="treelet restructuring rounds after every LBVH build (0: plain LBVH)")
    ap.add_argument("--trace-config", default=None, metavar="VC,RC,VA,RA[,SC,SA]",
                    help="fix the traversal variants, refill thresholds and slices")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the library copy kernel over BYTES (counter calibration)")
    args = ap.parse_args()
    args.pipeline = (PIPELINE.get(args.config) if args.pipeline == "auto" else
                     None if args.pipeline == "off" else args.pipeline)
    if args.treelets:
        CHILD_ARGS.extend(["--treelets", str(args.treelets)])
    if args.gpus != 1:
        print("bench.py measures one GPU: the ray-query path runs on a single MI355X (rays "
              "would split across GPUs with no exchange but a gather of the hits). "
              f"--gpus {args.gpus} is not supported.", file=sys.stderr)
        return 2
    ctx = rq.Context(0)
    if args.copy_calibration:
        nb = args.copy_calibration
        src, dst = rq.DeviceArray(ctx, nb), rq.DeviceArray(ctx, nb)
        for _ in range(3):
            ctx.copy(dst, src, nb)
        ctx.sync()
        print(json.dumps({"copy_bytes": nb, "launches": 3}), flush=True)
        return 0
    run = {"c3": run_c3, "c4": run_c4, "c5": run_c5}.get(args.config, run_frames)
    return run(ctx, args)


if __name__ == "__main__":
    sys.exit(main())
