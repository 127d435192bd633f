// Host side of the batched build + traversal (include/rq/rq_batch.h): per-mesh
// BVH buffers, the device work queue (items of every mesh, phase by phase),
// the persistent launch, its hipGraph capture and the batch autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "batch_internal.h"
#include "rq/rq_batch.h"
#include "rq_ctx.h"

using namespace rq;

namespace {
constexpr uint32_t kRayTiles[] = {256, 512, 1024, 2048, 8192};
constexpr uint32_t kRefills[] = {16, 32, 64};
constexpr uint32_t kMinRayTile = 256;
}  // namespace

struct rq_batch {
  std::vector<rq_batch_item> items;
  std::vector<rq_bvh*> bvhs;
  int mode = RQ_MORTON_SCENE_AABB;
  BatchMeshDev* meshes = nullptr;
  BatchWork* work = nullptr;
  size_t work_cap = 0;
  uint32_t* ctl = nullptr;
  int32_t* gstack = nullptr;
  uint32_t gdepth = 0;
  int build = 3;                       // build kernel: 5 waves per SIMD (batch_build_waves)
  std::vector<uint32_t> build_grid;    // its persistent grid, per build kernel
  std::vector<uint32_t> grid;          // the trace kernel's, per configuration
  // the autotuned choice on 512 mixed meshes (1k-500k triangles, incoherent
  // rays): trace configuration 1 (LDS depth 16), 512-ray items, refill at 16
  // idle lanes
  int variant = 1;
  uint32_t ray_tile = 512, refill = 16;
  uint32_t phase_begin[BATCH_PHASES + 1] = {};
  uint32_t trace_seg[BATCH_XCDS + 1] = {};
  uint32_t built_ray_tile = 0;
  hipStream_t own_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_variant = -1, graph_build = -1;
  uint32_t graph_ray_tile = 0, graph_refill = 0;
  std::vector<void*> allocs;
};

namespace {

int balloc(rq_ctx* c, rq_batch* b, size_t bytes, void** p) {
  int rc = ctx_alloc(c, bytes, p);
  if (rc == RQ_OK && *p) b->allocs.push_back(*p);
  return rc;
}

void drop_graph(rq_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_variant = -1;
}

void free_batch(rq_ctx* c, rq_batch* b) {
  drop_graph(b);
  for (rq_bvh* v : b->bvhs) {
    if (v->base) ctx_free(c, v->base);
    v->magic = 0;
    delete v;
  }
  for (void* p : b->allocs) ctx_free(c, p);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
}

// Trace items of mesh m: its closest-hit ray tiles, then its any-hit tiles.
void push_trace(std::vector<BatchWork>& w, const rq_batch_item& it, uint32_t m, uint32_t ray_tile) {
  const uint32_t nt = tiles_of(it.nrays, ray_tile);
  if (it.hits.prim)
    for (uint32_t t = 0; t < nt; ++t) w.push_back(BatchWork{m, t});
  if (it.occluded)
    for (uint32_t t = 0; t < nt; ++t) w.push_back(BatchWork{m, t | BATCH_ANY_BIT});
}

// The work queue: the items of every mesh, phase after phase, tiles in order.
// Trace items go to BATCH_XCDS segments: meshes are dealt to the segment with
// the least ray work so far, largest first (LPT), and keep their index order
// inside a segment.
int make_work(rq_ctx* c, rq_batch* b, uint32_t ray_tile) {
  // a captured graph holds this queue's phase bounds in its kernel arguments
  drop_graph(b);
  std::vector<BatchWork> w;
  const uint32_t nm = (uint32_t)b->items.size();
  for (int ph = 0; ph < BATCH_PHASES; ++ph) {
    b->phase_begin[ph] = (uint32_t)w.size();
    if (ph == PH_TRACE) {
      std::vector<uint64_t> work(nm);
      std::vector<uint32_t> order(nm), seg(nm);
      for (uint32_t m = 0; m < nm; ++m) {
        const rq_batch_item& it = b->items[m];
        work[m] = (uint64_t)it.nrays * ((it.hits.prim ? 1 : 0) + (it.occluded ? 1 : 0));
        order[m] = m;
      }
      std::stable_sort(order.begin(), order.end(),
                       [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
      uint64_t load[BATCH_XCDS] = {};
      for (uint32_t m : order) {
        const int q = (int)(std::min_element(load, load + BATCH_XCDS) - load);
        seg[m] = (uint32_t)q;
        load[q] += work[m];
      }
      for (int q = 0; q < BATCH_XCDS; ++q) {
        b->trace_seg[q] = (uint32_t)w.size();
        for (uint32_t m = 0; m < nm; ++m)
          if (seg[m] == (uint32_t)q) push_trace(w, b->items[m], m, ray_tile);
      }
      b->trace_seg[BATCH_XCDS] = (uint32_t)w.size();
      continue;
    }
    for (uint32_t m = 0; m < nm; ++m) {
      const uint32_t n = b->bvhs[m]->n;
      uint32_t k = 0;
      switch (ph) {
        case PH_INIT: k = n ? 1 : 0; break;
        case PH_BOUNDS: k = (n && b->mode == RQ_MORTON_SCENE_AABB) ? tiles_of(n, TRI_TILE) : 0; break;
        case PH_MORTON: k = tiles_of(n, TRI_TILE); break;
        case PH_SORT0: case PH_SORT1: case PH_SORT2: case PH_SORT3:
          k = tiles_of(n, SORT_TILE_LEAN);
          break;
        case PH_KARRAS: k = n >= 2 ? tiles_of(n - 1, NODE_TILE) : 0; break;
        case PH_REFIT: k = tiles_of(n, BUILD_BLOCK * BATCH_REFIT_CHUNKS); break;
        // spanning nodes exist once there are two chunks; pass 1 writes the
        // wide records of all the others
        case PH_TOP:
        case PH_WIDEN:
          k = tiles_of(n, BUILD_BLOCK) > 1 ? tiles_of(n - 1, BUILD_BLOCK * BATCH_SPAN_TILES) : 0;
          break;
        default: break;
      }
      for (uint32_t t = 0; t < k; ++t) w.push_back(BatchWork{m, t});
    }
  }
  b->phase_begin[BATCH_PHASES] = (uint32_t)w.size();
  if (w.size() > b->work_cap) return set_err(c, RQ_E_INVALID, "work queue larger than reserved");
  if (!w.empty()) {
    hipError_t e = hipMemcpy(b->work, w.data(), w.size() * sizeof(BatchWork), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err(c, e, "work queue upload");
  }
  b->built_ray_tile = ray_tile;
  return RQ_OK;
}

BatchArgs batch_args(const rq_batch* b) {
  BatchArgs a{};
  a.meshes = b->meshes;
  a.items = b->work;
  std::memcpy(a.phase_begin, b->phase_begin, sizeof a.phase_begin);
  std::memcpy(a.trace_seg, b->trace_seg, sizeof a.trace_seg);
  a.ctl = b->ctl;
  a.gstack = b->gstack;
  int depth, ts;
  batch_variant_info(b->variant, &depth, &ts);
  a.gdepth = TRACE_MAX_STACK_WIDE - depth;  // the batch run walks the 4-wide nodes
  a.ray_tile = b->ray_tile;
  a.refill_idle = b->refill;
  return a;
}

int launch(rq_ctx* c, rq_batch* b, hipStream_t s, bool profile) {
  if (b->built_ray_tile != b->ray_tile) {
    int rc = make_work(c, b, b->ray_tile);
    if (rc) return rc;
  }
  LaunchCtx lc = launch_ctx(c, s);
  if (!profile) lc.prof = nullptr;
  hipError_t e =
      launch_batch(lc, batch_args(b), b->build, b->build_grid[b->build], b->variant,
                   b->grid[b->variant]);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "batch launch");
}

}  // namespace

extern "C" {

int rq_batch_create(rq_ctx* c, const rq_batch_item* items, uint32_t count, int mode,
                    rq_batch** out) {
  if (!c || !out || (count && !items)) return RQ_E_INVALID;
  *out = nullptr;
  if (mode != RQ_MORTON_REF_AFFINE && mode != RQ_MORTON_SCENE_AABB)
    return set_err(c, RQ_E_INVALID, "unknown morton mode %d", mode);
  for (uint32_t i = 0; i < count; ++i) {
    const rq_batch_item& it = items[i];
    if (!it.scene || it.scene->magic != SCENE_MAGIC)
      return set_err(c, RQ_E_INVALID, "item %u: bad scene", i);
    if (it.scene->kind != RQ_PRIM_TRIANGLE)
      return set_err(c, RQ_E_UNSUPPORTED, "item %u: batches hold triangle scenes", i);
    if (it.nrays && (!it.rays.ox || !it.rays.oy || !it.rays.oz || !it.rays.dx || !it.rays.dy ||
                     !it.rays.dz || !it.rays.tmin || !it.rays.tmax))
      return set_err(c, RQ_E_INVALID, "item %u: null ray array", i);
  }
  (void)hipSetDevice(c->device);
  auto* b = new rq_batch();
  b->items.assign(items, items + count);
  b->mode = mode;
  int rc = RQ_OK;
  for (uint32_t i = 0; i < count && !rc; ++i) {
    auto* v = new rq_bvh();
    v->magic = BVH_MAGIC;
    v->scene = items[i].scene;
    v->mode = mode;
    v->n = items[i].scene->ntri;
    v->owned = false;
    b->bvhs.push_back(v);
    rc = bvh_alloc(c, v, v->n);
  }
  // device mesh table
  std::vector<BatchMeshDev> md(count);
  size_t work_cap = 0;
  for (uint32_t i = 0; i < count && !rc; ++i) {
    const rq_batch_item& it = items[i];
    const rq_bvh* v = b->bvhs[i];
    const rq_scene* s = it.scene;
    BatchMeshDev& d = md[i];
    d.v = BuildView{s->vx, s->vy, s->vz, s->idx, v->n, mode, v->B, s->r, s->r != nullptr};
    d.rays = it.rays;
    d.nrays = it.nrays;
    d.prim = it.hits.prim, d.t = it.hits.t, d.u = it.hits.u, d.v_ = it.hits.v;
    d.occ = it.occluded;
    const uint32_t n = v->n;
    // init, bounds + Morton, sort, Karras, refit, spanning nodes + their wide
    // records, trace
    work_cap += 1 + 2 * tiles_of(n, TRI_TILE) + SORT_PASSES * tiles_of(n, SORT_TILE_LEAN) +
                tiles_of(n, NODE_TILE) + tiles_of(n, BUILD_BLOCK) +
                2 * (size_t)tiles_of(n, BUILD_BLOCK) + 2 * (size_t)tiles_of(it.nrays, kMinRayTile);
  }
  if (!rc) rc = balloc(c, b, std::max<size_t>(1, count) * sizeof(BatchMeshDev), (void**)&b->meshes);
  if (!rc && count) {
    hipError_t e = hipMemcpy(b->meshes, md.data(), count * sizeof(BatchMeshDev),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_err(c, e, "batch mesh table");
  }
  if (!rc) {
    b->work_cap = std::max<size_t>(work_cap, 1);
    rc = balloc(c, b, b->work_cap * sizeof(BatchWork), (void**)&b->work);
  }
  if (!rc) rc = balloc(c, b, BATCH_CTL_ALLOC_WORDS * sizeof(uint32_t), (void**)&b->ctl);
  if (!rc) {
    hipError_t e = hipMemset(b->ctl, 0, BATCH_CTL_ALLOC_WORDS * sizeof(uint32_t));
    if (e != hipSuccess) rc = hip_err(c, e, "batch control init");
  }
  if (!rc) {
    size_t gmax = 0;
    for (int k = 0; k < batch_num_build_variants(); ++k)
      b->build_grid.push_back(batch_build_grid_blocks(c->device, k));
    for (int v = 0; v < batch_num_variants(); ++v) {
      b->grid.push_back(batch_trace_grid_blocks(c->device, v));
      int depth, ts;
      batch_variant_info(v, &depth, &ts);
      gmax = std::max(gmax,
                      (size_t)b->grid.back() * BUILD_BLOCK * (TRACE_MAX_STACK_WIDE - depth));
    }
    rc = balloc(c, b, gmax * 2 * sizeof(int32_t), (void**)&b->gstack);
  }
  if (!rc) rc = make_work(c, b, b->ray_tile);
  if (!rc) {
    hipError_t e = hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = hip_err(c, e, "batch stream");
  }
  if (rc) {
    free_batch(c, b);
    return rc;
  }
  *out = b;
  return RQ_OK;
}

int rq_batch_destroy(rq_ctx* c, rq_batch* b) {
  if (!c || !b) return RQ_E_INVALID;
  (void)hipDeviceSynchronize();
  free_batch(c, b);
  return RQ_OK;
}

int rq_batch_run(rq_ctx* c, rq_batch* b, void* stream) {
  if (!c || !b) return RQ_E_INVALID;
  return launch(c, b, (hipStream_t)stream, true);
}

int rq_batch_run_graph(rq_ctx* c, rq_batch* b, void* stream) {
  if (!c || !b) return RQ_E_INVALID;
  // the replay goes on the caller's stream (the null stream included), in
  // order after the caller's earlier work there, rq_batch_run's among it:
  // two runs of one batch share its control words and must not overlap
  hipStream_t s = (hipStream_t)stream;
  if (b->built_ray_tile != b->ray_tile) {
    int rc = make_work(c, b, b->ray_tile);  // also drops the graph
    if (rc) return rc;
  }
  if (!b->exec || b->graph_variant != b->variant || b->graph_build != b->build ||
      b->graph_ray_tile != b->ray_tile || b->graph_refill != b->refill) {
    drop_graph(b);
    hipError_t e = hipStreamBeginCapture(b->own_stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return hip_err(c, e, "graph capture begin");
    int rc = launch(c, b, b->own_stream, false);
    e = hipStreamEndCapture(b->own_stream, &b->graph);
    if (rc) return rc;
    if (e != hipSuccess) return hip_err(c, e, "graph capture end");
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return hip_err(c, e, "graph instantiate");
    }
    b->graph_variant = b->variant;
    b->graph_build = b->build;
    b->graph_ray_tile = b->ray_tile;
    b->graph_refill = b->refill;
  }
  LaunchCtx lc = launch_ctx(c, s);
  ProfScope ps(lc, "batch_graph");
  hipError_t e = hipGraphLaunch(b->exec, s);
  if (e == hipSuccess && !stream) e = hipStreamSynchronize(s);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "graph launch");
}

int rq_batch_bvh(rq_ctx* c, rq_batch* b, uint32_t i, rq_bvh** out) {
  if (!c || !b || !out || i >= b->bvhs.size()) return RQ_E_INVALID;
  *out = b->bvhs[i];
  return RQ_OK;
}

int rq_batch_config_count(void) { return batch_num_variants(); }

int rq_batch_config_get(int index, rq_batch_config* out) {
  if (index < 0 || index >= batch_num_variants() || !out) return RQ_E_INVALID;
  int depth, ts;
  batch_variant_info(index, &depth, &ts);
  out->lds_depth = depth;
  out->tstack = ts;
  out->ray_tile = 0;
  out->refill_idle = 0;
  return RQ_OK;
}

int rq_batch_set_config(rq_ctx* c, rq_batch* b, int index, uint32_t ray_tile,
                        uint32_t refill_idle) {
  if (!c || !b) return RQ_E_INVALID;
  if (index < 0 || index >= batch_num_variants() || ray_tile < kMinRayTile || refill_idle < 1 ||
      refill_idle > 64)
    return set_err(c, RQ_E_INVALID, "bad batch configuration");
  b->variant = index;
  b->ray_tile = ray_tile;
  b->refill = refill_idle;
  return RQ_OK;
}

int rq_batch_autotune(rq_ctx* c, rq_batch* b, int repeats, int* best_index,
                      uint32_t* best_ray_tile, uint32_t* best_refill, double* best_ms) {
  if (!c || !b || repeats < 1) return RQ_E_INVALID;
  // the timed runs go on the batch's own stream: after every earlier run
  hipError_t se = hipDeviceSynchronize();
  if (se != hipSuccess) return hip_err(c, se, "batch autotune");
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double best = 1e30;
  int bv = b->variant;
  uint32_t bt = b->ray_tile, br = b->refill;
  int rc = RQ_OK;
  for (int v = 0; v < batch_num_variants() && !rc; ++v)
    for (uint32_t tile : kRayTiles)
      for (uint32_t rf : kRefills) {
        b->variant = v, b->ray_tile = tile, b->refill = rf;
        for (int k = 0; k <= repeats && !rc; ++k) {
          if (k == 1) (void)hipEventRecord(e0, b->own_stream);
          rc = launch(c, b, b->own_stream, false);
        }
        (void)hipEventRecord(e1, b->own_stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (!rc && ms / repeats < best) best = ms / repeats, bv = v, bt = tile, br = rf;
      }
  // then the build kernel, with the chosen trace configuration
  b->variant = bv, b->ray_tile = bt, b->refill = br;
  int bb = b->build;
  for (int k = 0; k < batch_num_build_variants() && !rc; ++k) {
    b->build = k;
    for (int r = 0; r <= repeats && !rc; ++r) {
      if (r == 1) (void)hipEventRecord(e0, b->own_stream);
      rc = launch(c, b, b->own_stream, false);
    }
    (void)hipEventRecord(e1, b->own_stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (!rc && ms / repeats < best) best = ms / repeats, bb = k;
  }
  b->build = bb;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (best_index) *best_index = bv;
  if (best_ray_tile) *best_ray_tile = bt;
  if (best_refill) *best_refill = br;
  if (best_ms) *best_ms = best;
  return RQ_OK;
}

int rq_batch_phase_items(rq_ctx* c, rq_batch* b, uint32_t* counts, int cap) {
  if (!c || !b || !counts) return RQ_E_INVALID;
  for (int ph = 0; ph < BATCH_PHASES && ph < cap; ++ph)
    counts[ph] = b->phase_begin[ph + 1] - b->phase_begin[ph];
  return BATCH_PHASES;
}

int rq_batch_phase_ms(rq_ctx* c, rq_batch* b, double* ms, int cap) {
  if (!c || !b || !ms) return RQ_E_INVALID;
  uint32_t w[BATCH_CTL_WORDS];
  int khz = 0;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(w, b->ctl, sizeof w, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device);
  if (e != hipSuccess) return hip_err(c, e, "batch phase times");
  if (khz <= 0) return set_err(c, RQ_E_HIP, "no wall-clock rate");
  uint64_t st[BATCH_PHASES];
  std::memcpy(st, w + BATCH_CTL_CLOCK, sizeof st);
  for (int ph = 0; ph < (int)PH_TRACE && ph < cap; ++ph)
    ms[ph] = st[ph + 1] >= st[ph] && st[0] != ~0ull ? (double)(st[ph + 1] - st[ph]) / khz : -1.0;
  return (int)PH_TRACE;
}

int rq_batch_errors(rq_ctx* c, rq_batch* b, uint32_t* bits) {
  if (!c || !b || !bits) return RQ_E_INVALID;
  uint32_t w[BATCH_CTL_WORDS];
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(w, b->ctl, sizeof w, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "batch errors");
  *bits = (w[BATCH_CTL_BARRIER_ERR] ? 1u : 0u) | (w[BATCH_CTL_TRACE_ERR] ? 2u : 0u);
  for (rq_bvh* v : b->bvhs) {
    if (!v->n) continue;
    uint32_t m = 0;
    e = hipMemcpy(&m, v->B.misc + MISC_ERR, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err(c, e, "batch errors");
    if (m) *bits |= 4u;
  }
  return RQ_OK;
}

int rq_batch_get_config(rq_ctx* c, const rq_batch* b, int* index, uint32_t* ray_tile,
                        uint32_t* refill_idle) {
  if (!c || !b || !index || !ray_tile || !refill_idle) return RQ_E_INVALID;
  *index = b->variant, *ray_tile = b->ray_tile, *refill_idle = b->refill;
  return RQ_OK;
}

int rq_batch_set_build_waves(rq_ctx* c, rq_batch* b, int waves) {
  if (!c || !b) return RQ_E_INVALID;
  for (int k = 0; k < batch_num_build_variants(); ++k)
    if (batch_build_waves(k) == waves) {
      b->build = k;
      return RQ_OK;
    }
  return set_err(c, RQ_E_INVALID, "no build kernel with %d waves per SIMD", waves);
}

int rq_batch_get_build_waves(rq_ctx* c, const rq_batch* b, int* waves) {
  if (!c || !b || !waves) return RQ_E_INVALID;
  *waves = batch_build_waves(b->build);
  return RQ_OK;
}

int rq_batch_occupancy(rq_ctx* c, rq_batch* b, uint32_t* members, uint32_t* grid) {
  if (!c || !b || !members || !grid) return RQ_E_INVALID;
  uint32_t m = 0;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(&m, b->ctl + BATCH_CTL_MEMBERS, sizeof m, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "batch occupancy");
  *members = m;
  *grid = b->build_grid[b->build];
  return RQ_OK;
}

}  // extern "C"
