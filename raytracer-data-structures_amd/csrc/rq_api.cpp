// C ABI of librq (include/rq/rq.h): contexts, HBM budget accounting, scene
// upload, LBVH build/rebuild/export, closest/any-hit queries, per-kernel
// event timing. Host code only; kernels live in the .hip files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "rq_ctx.h"

namespace rq {

struct Profiler {
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, uint64_t>> totals;
  std::set<std::string> only;  // kernel names to time (empty: every kernel)
  void resolve() {
    for (auto& p : pending) {
      float ms = 0.0f;
      hipEventSynchronize(p.b);
      hipEventElapsedTime(&ms, p.a, p.b);
      auto& t = totals[p.name];
      t.first += ms;
      t.second += 1;
      hipEventDestroy(p.a);
      hipEventDestroy(p.b);
    }
    pending.clear();
  }
};

void prof_begin(Profiler* p, hipStream_t s, const char* name, hipEvent_t* ev) {
  if (!p->only.empty() && !p->only.count(name)) {
    *ev = nullptr;  // not timed: no events on the stream
    return;
  }
  hipEventCreate(ev);
  hipEventRecord(*ev, s);
}
void prof_end(Profiler* p, hipStream_t s, const char* name, hipEvent_t ev) {
  if (!ev) return;
  hipEvent_t b;
  hipEventCreate(&b);
  hipEventRecord(b, s);
  p->pending.push_back({name, ev, b});
  if (p->pending.size() > 4096) p->resolve();
}

int set_err(rq_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

int hip_err(rq_ctx* c, hipError_t e, const char* what) {
  return set_err(c, RQ_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int ctx_alloc(rq_ctx* c, size_t bytes, void** p) {
  *p = nullptr;
  if (bytes == 0) return RQ_OK;
  if (c->used + bytes > c->budget)
    return set_err(c, RQ_E_NOMEM, "HBM budget exceeded: %zu in use + %zu requested > %zu budget",
                   c->used, bytes, c->budget);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(c, RQ_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  c->used += bytes;
  c->sizes[*p] = bytes;
  return RQ_OK;
}

void ctx_free(rq_ctx* c, void* p) {
  if (!p) return;
  auto it = c->sizes.find(p);
  if (it != c->sizes.end()) {
    c->used -= it->second;
    c->sizes.erase(it);
  }
  hipFree(p);
}

LaunchCtx launch_ctx(rq_ctx* c, void* stream) {
  return LaunchCtx{(hipStream_t)stream, c->prof_on ? c->prof : nullptr};
}

uint32_t next_counter_slot(rq_ctx* c) { return c->counter_seq++ % COUNTER_SLOTS; }

int32_t* stream_gstack(rq_ctx* c, void* stream, bool may_alloc) {
  if (!stream) return c->gstack;
  for (auto& kv : c->gstack_streams)
    if (kv.first == stream) return kv.second;
  if (!may_alloc) {
    set_err(c, RQ_E_UNSUPPORTED,
            "first traversal on a stream under graph capture: run one query on the stream "
            "before capturing (its stack overflow area is allocated then)");
    return nullptr;
  }
  if (c->gstack_streams.size() >= MAX_TRACE_STREAMS) {
    // take over the area of a stream whose work has all completed (a stream
    // destroyed without rq_stream_forget reports an error: also free)
    for (auto& kv : c->gstack_streams)
      if (hipStreamQuery((hipStream_t)kv.first) != hipErrorNotReady) {
        kv.first = stream;
        return kv.second;
      }
    set_err(c, RQ_E_NOMEM, "traversal launches on more than %d busy streams",
            MAX_TRACE_STREAMS);
    return nullptr;
  }
  void* p = nullptr;
  if (ctx_alloc(c, c->gstack_bytes, &p) != RQ_OK) return nullptr;
  c->gstack_streams.push_back({stream, static_cast<int32_t*>(p)});
  return static_cast<int32_t*>(p);
}

}  // namespace rq

using namespace rq;

namespace rq {

int ensure_height(rq_ctx* c, const rq_bvh* b) {
  if (b->n < 2) return RQ_OK;
  uint32_t h = 0;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(&h, b->B.misc + MISC_HEIGHT, 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && h == HEIGHT_UNKNOWN) {
    BuildBuffers B = b->B;
    e = launch_height(LaunchCtx{nullptr, nullptr}, B, b->n);
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "tree height");
}

int bvh_alloc(rq_ctx* c, rq_bvh* b, uint32_t n) {
  BuildBuffers& B = b->B;
  B.ntiles = (n + SORT_TILE_LEAN - 1) / SORT_TILE_LEAN;
  size_t nn = std::max<uint32_t>(n, 1);
  size_t ni = n > 1 ? n - 1 : 1;
  size_t sz[] = {4 * nn,
                 4 * nn,
                 4 * nn,
                 4 * nn,
                 4 * nn,
                 (size_t)SORT_PASSES * std::max<uint32_t>(B.ntiles, 1) * 256 * 4,
                 (size_t)MISC_WORDS * 4,
                 4 * (2 * nn),
                 4 * ni,
                 8 * ni,
                 8 * ni,
                 64 * ni,
                 48 * nn,
                 4 * nn,
                 128 * ni,
                 24 * nn,
                 24 * nn,
                 24 * (size_t)((nn + REFIT_CHUNK_MIN - 1) / REFIT_CHUNK_MIN)};
  size_t total = 0;
  for (size_t s : sz) total += align256(s);
  void* p;
  int rc = ctx_alloc(c, total, &p);
  if (rc) return rc;
  char* q = (char*)p;
  void** dst[] = {(void**)&B.codes, (void**)&B.keys_a, (void**)&B.keys_b, (void**)&B.vals_a,
                  (void**)&B.vals_b, (void**)&B.status, (void**)&B.misc, (void**)&B.parent,
                  (void**)&B.flags, (void**)&B.slot_h, (void**)&B.range, (void**)&B.nodes,
                  (void**)&B.tris, (void**)&B.frontier, (void**)&B.wide,
                  (void**)&B.pre, (void**)&B.suf, (void**)&B.cbox};
  for (size_t i = 0; i < sizeof(sz) / sizeof(sz[0]); ++i) {
    *dst[i] = sz[i] ? q : nullptr;
    q += align256(sz[i]);
  }
  b->base = p;
  b->bytes = total;
  return RQ_OK;
}

}  // namespace rq

namespace {
int check_ctx(rq_ctx* c) { return c ? RQ_OK : RQ_E_INVALID; }
}  // namespace

extern "C" {

const char* rq_version(void) { return "rq 0.1 (gfx950)"; }

int rq_create(rq_ctx** out, int device, size_t budget) {
  if (!out) return RQ_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RQ_E_NODEVICE;
  if (device < 0 || device >= ndev) return RQ_E_INVALID;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return RQ_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return RQ_E_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return RQ_E_NODEVICE;
  auto* c = new rq_ctx();
  c->device = device;
  c->budget = budget ? budget : (size_t)prop.totalGlobalMem;
  c->prof = new Profiler();
  size_t gmax = 0;
  for (int v = 0; v < trace_num_variants(); ++v) {
    c->trace_grid[v] = trace_grid_blocks(device, v, false);
    c->trace_grid_any[v] = trace_grid_blocks(device, v, true);
    int blk, depth, ts;
    trace_variant_info(v, &blk, &depth, &ts);
    gmax = std::max(gmax, (size_t)std::max(c->trace_grid[v], c->trace_grid_any[v]) * blk *
                              (trace_max_stack(v) - depth));
  }
  // counter slots (eager, then capture) + one trailing word of traversal error bits
  const size_t cbytes = ((COUNTER_SLOTS + CAPTURE_SLOTS) * 2 + 1) * sizeof(uint32_t);
  int rc = ctx_alloc(c, cbytes, (void**)&c->counters);
  if (rc == RQ_OK) {
    c->errword = c->counters + (COUNTER_SLOTS + CAPTURE_SLOTS) * 2;
    hipMemset(c->counters, 0, cbytes);
    c->gstack_bytes = gmax * 2 * sizeof(int32_t);
    rc = ctx_alloc(c, c->gstack_bytes, (void**)&c->gstack);
  }
  if (rc == RQ_OK) {
    void* sp = nullptr;
    const hipError_t e = noalloc_scratch_address(&sp);
    if (e != hipSuccess || !sp) rc = hip_err(c, e, "scratch symbol address");
    c->scratch = static_cast<unsigned char*>(sp);
  }
  if (rc == RQ_OK && hipDeviceSynchronize() != hipSuccess) rc = RQ_E_HIP;
  if (rc) {
    rq_destroy(c);
    return rc;
  }
  *out = c;
  return RQ_OK;
}

int rq_destroy(rq_ctx* c) {
  if (!c) return RQ_E_INVALID;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  if (c->prof) c->prof->resolve();
  for (hipEvent_t e : c->order_events) (void)hipEventDestroy(e);
  stage_release(c);
  for (auto& kv : c->sizes) hipFree(kv.first);
  c->sizes.clear();
  delete c->prof;
  delete c;
  return RQ_OK;
}

const char* rq_last_error(const rq_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t rq_bytes_in_use(const rq_ctx* c) { return c ? c->used : 0; }

int rq_scene_upload(rq_ctx* c, const float* vx, const float* vy, const float* vz, uint32_t nv,
                    const uint32_t* idx, uint32_t ntri, int prim_kind, rq_scene** out,
                    void* stream) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (!out) return set_err(c, RQ_E_INVALID, "null output handle");
  *out = nullptr;
  if (prim_kind != RQ_PRIM_TRIANGLE)
    return set_err(c, RQ_E_UNSUPPORTED, "primitive kind %d not supported", prim_kind);
  if (ntri > MAX_PRIMS) return set_err(c, RQ_E_INVALID, "too many triangles (%u)", ntri);
  if ((nv && (!vx || !vy || !vz)) || (ntri && !idx))
    return set_err(c, RQ_E_INVALID, "null vertex or index array");
  for (uint64_t k = 0; k < 3ull * ntri; ++k)
    if (idx[k] >= nv) return set_err(c, RQ_E_INVALID, "index %u out of range (nv=%u)", idx[k], nv);
  hipSetDevice(c->device);
  auto* s = new rq_scene();
  s->magic = SCENE_MAGIC;
  s->nv = nv;
  s->ntri = ntri;
  size_t vb = align256((size_t)nv * 4), ib = align256((size_t)ntri * 12);
  void* p;
  int rc = ctx_alloc(c, 3 * vb + ib, &p);
  if (rc) {
    delete s;
    return rc;
  }
  s->base = p;
  s->bytes = 3 * vb + ib;
  s->vx = (float*)p;
  s->vy = (float*)((char*)p + vb);
  s->vz = (float*)((char*)p + 2 * vb);
  s->idx = (uint32_t*)((char*)p + 3 * vb);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (nv) {
    e = hipMemcpyAsync(s->vx, vx, (size_t)nv * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->vy, vy, (size_t)nv * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->vz, vz, (size_t)nv * 4, hipMemcpyHostToDevice, st);
  }
  if (e == hipSuccess && ntri)
    e = hipMemcpyAsync(s->idx, idx, (size_t)ntri * 12, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // host arrays may go away after return
  if (e != hipSuccess) {
    ctx_free(c, p);
    delete s;
    return hip_err(c, e, "scene upload");
  }
  *out = s;
  return RQ_OK;
}

int rq_scene_upload_spheres(rq_ctx* c, const float* cx, const float* cy, const float* cz,
                            const float* radius, const float* rgb, uint32_t n, rq_scene** out,
                            void* stream) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (!out) return set_err(c, RQ_E_INVALID, "null output handle");
  *out = nullptr;
  if (n > MAX_PRIMS) return set_err(c, RQ_E_INVALID, "too many spheres (%u)", n);
  if (n && (!cx || !cy || !cz || !radius)) return set_err(c, RQ_E_INVALID, "null sphere array");
  hipSetDevice(c->device);
  auto* s = new rq_scene();
  s->magic = SCENE_MAGIC;
  s->kind = RQ_PRIM_SPHERE;
  s->nv = s->ntri = n;
  const size_t fb = align256((size_t)n * 4), cb = rgb ? align256((size_t)n * 12) : 0;
  void* p;
  int rc = ctx_alloc(c, 4 * fb + cb, &p);
  if (rc) {
    delete s;
    return rc;
  }
  s->base = p;
  s->bytes = 4 * fb + cb;
  char* q = (char*)p;
  s->vx = (float*)q, s->vy = (float*)(q + fb), s->vz = (float*)(q + 2 * fb);
  s->r = (float*)(q + 3 * fb);
  if (rgb) s->rgb = (float*)(q + 4 * fb);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (n) {
    e = hipMemcpyAsync(s->vx, cx, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->vy, cy, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->vz, cz, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->r, radius, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && rgb)
      e = hipMemcpyAsync(s->rgb, rgb, (size_t)n * 12, hipMemcpyHostToDevice, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // host arrays may go away after return
  if (e != hipSuccess) {
    ctx_free(c, p);
    delete s;
    return hip_err(c, e, "sphere upload");
  }
  *out = s;
  return RQ_OK;
}

int rq_scene_set_materials(rq_ctx* c, rq_scene* s, const uint8_t* mat, void* stream) {
  if (check_ctx(c) || !s || s->magic != SCENE_MAGIC || !mat)
    return set_err(c, RQ_E_INVALID, "bad argument");
  for (uint32_t i = 0; i < s->ntri; ++i)
    if (mat[i] > RQ_MAT_REFLECTION) return set_err(c, RQ_E_INVALID, "material %u at %u", mat[i], i);
  hipSetDevice(c->device);
  if (!s->mat) {
    int rc = ctx_alloc(c, std::max<size_t>(s->ntri, 1), (void**)&s->mat);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = s->ntri ? hipMemcpyAsync(s->mat, mat, s->ntri, hipMemcpyHostToDevice, st)
                         : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "material upload");
}

int rq_scene_update_vertices(rq_ctx* c, rq_scene* s, const float* vx, const float* vy,
                             const float* vz, void* stream) {
  if (check_ctx(c) || !s || s->magic != SCENE_MAGIC) return set_err(c, RQ_E_INVALID, "bad scene");
  hipStream_t st = (hipStream_t)stream;
  size_t b = (size_t)s->nv * 4;
  hipError_t e = hipMemcpyAsync(s->vx, vx, b, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->vy, vy, b, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->vz, vz, b, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "vertex update");
}

int rq_build_lbvh(rq_ctx* c, const rq_scene* s, int mode, rq_bvh** out, void* stream) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (!out) return set_err(c, RQ_E_INVALID, "null output handle");
  *out = nullptr;
  if (!s || s->magic != SCENE_MAGIC) return set_err(c, RQ_E_INVALID, "bad scene handle");
  if (mode != RQ_MORTON_REF_AFFINE && mode != RQ_MORTON_SCENE_AABB)
    return set_err(c, RQ_E_INVALID, "unknown morton mode %d", mode);
  hipSetDevice(c->device);
  auto* b = new rq_bvh();
  b->magic = BVH_MAGIC;
  b->scene = s;
  b->mode = mode;
  b->n = s->ntri;
  int rc = bvh_alloc(c, b, b->n);
  if (rc) {
    delete b;
    return rc;
  }
  rc = rq_rebuild_lbvh(c, b, stream);
  if (rc) {
    ctx_free(c, b->base);
    delete b;
    return rc;
  }
  *out = b;
  return RQ_OK;
}

int rq_rebuild_lbvh(rq_ctx* c, rq_bvh* b, void* stream) {
  if (check_ctx(c) || !b || b->magic != BVH_MAGIC) return set_err(c, RQ_E_INVALID, "bad bvh");
  const rq_scene* s = b->scene;
  hipError_t e = launch_build(launch_ctx(c, stream), s->vx, s->vy, s->vz, s->idx, b->n, b->mode,
                              b->B, s->r);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "lbvh build launch");
}

int rq_bvh_treelets(rq_ctx* c, rq_bvh* b, int rounds, void* stream) {
  if (check_ctx(c) || !b || b->magic != BVH_MAGIC) return set_err(c, RQ_E_INVALID, "bad bvh");
  if (rounds < 1 || rounds > 8) return set_err(c, RQ_E_INVALID, "rounds outside 1..8");
  if (b->n < 7) return RQ_OK;
  int rc = ensure_height(c, b);
  if (rc) return rc;
  uint32_t misc[MISC_WORDS];
  hipError_t e = hipMemcpy(misc, b->B.misc, sizeof misc, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "treelets");
  if (misc[MISC_ERR]) return set_err(c, RQ_E_INVALID, "the BVH's build reported errors");
  if (misc[MISC_HEIGHT] > (uint32_t)TRACE_MAX_STACK)
    return set_err(c, RQ_E_INVALID, "tree height %u out of range", misc[MISC_HEIGHT]);
  const LaunchCtx lc = launch_ctx(c, stream);
  e = launch_treelets(lc, b->B, b->n, misc[MISC_HEIGHT], rounds);
  if (e == hipSuccess && b->B.wide) e = launch_widen(lc, b->B, b->n);
  if (e == hipSuccess) e = launch_height(lc, b->B, b->n);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "treelets launch");
}

int rq_bvh_info_get(rq_ctx* c, const rq_bvh* b, rq_bvh_info* info) {
  if (check_ctx(c) || !b || b->magic != BVH_MAGIC || !info)
    return set_err(c, RQ_E_INVALID, "bad argument");
  std::memset(info, 0, sizeof *info);
  info->n_prims = b->n;
  info->n_internal = b->n > 1 ? b->n - 1 : 0;
  info->morton_mode = b->mode;
  info->device_bytes = b->bytes;
  if (b->n == 0) return RQ_OK;
  int rc = ensure_height(c, b);
  if (rc) return rc;
  uint32_t misc[MISC_WORDS];
  hipError_t e = hipMemcpy(misc, b->B.misc, sizeof misc, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "bvh info");
  std::memcpy(info->root_lo, misc + MISC_ROOT, 12);
  std::memcpy(info->root_hi, misc + MISC_ROOT + 3, 12);
  info->height = misc[MISC_HEIGHT];
  if (misc[MISC_ERR]) return set_err(c, RQ_E_HIP, "build reported error bits 0x%x", misc[MISC_ERR]);
  return RQ_OK;
}

int rq_bvh_export(rq_ctx* c, const rq_bvh* b, uint32_t* codes, uint32_t* sorted_idx,
                  int32_t* left, int32_t* right, float* boxes) {
  if (check_ctx(c) || !b || b->magic != BVH_MAGIC) return set_err(c, RQ_E_INVALID, "bad bvh");
  const uint32_t n = b->n;
  if (n == 0) return RQ_OK;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_err(c, e, "export sync");
  if (codes) e = hipMemcpy(codes, b->B.codes, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && sorted_idx)
    e = hipMemcpy(sorted_idx, b->B.vals_a, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "export copy");
  if (!left && !right && !boxes) return RQ_OK;
  const uint32_t ni = n - 1;
  std::vector<NodeRec> nodes(ni);
  std::vector<int32_t> parent(2 * (size_t)n - 1);
  std::vector<float> misc(MISC_WORDS);
  if (ni) e = hipMemcpy(nodes.data(), b->B.nodes, (size_t)ni * 64, hipMemcpyDeviceToHost);
  if (e == hipSuccess && ni)
    e = hipMemcpy(parent.data(), b->B.parent, parent.size() * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(misc.data(), b->B.misc, MISC_WORDS * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "export copy");
  uint32_t ebits;
  std::memcpy(&ebits, misc.data() + MISC_ERR, 4);
  if (ebits) return set_err(c, RQ_E_HIP, "build reported error bits 0x%x", ebits);
  for (uint32_t i = 0; i < ni; ++i) {
    int32_t l, r;
    std::memcpy(&l, &nodes[i].lo0.w, 4);
    std::memcpy(&r, &nodes[i].hi0.w, 4);
    if (left) left[i] = l;
    if (right) right[i] = r;
  }
  if (boxes) {
    auto slot = [&](uint32_t node, uint32_t side, float* out6) {
      const NodeRec& q = nodes[node];
      const float4& lo = side ? q.lo1 : q.lo0;
      const float4& hi = side ? q.hi1 : q.hi0;
      out6[0] = lo.x, out6[1] = lo.y, out6[2] = lo.z, out6[3] = hi.x, out6[4] = hi.y,
      out6[5] = hi.z;
    };
    // every node except the root has its box in its parent's child slot
    for (uint32_t u = 0; u < 2 * n - 1; ++u) {
      float* out6 = boxes + 6 * (size_t)u;
      if (u == 0 && n > 1) {
        std::memcpy(out6, misc.data() + MISC_ROOT, 24);
      } else if (n == 1) {
        std::memcpy(out6, misc.data() + MISC_ROOT, 24);
      } else {
        int32_t pe = parent[u];
        slot((uint32_t)pe >> 1, (uint32_t)pe & 1u, out6);
      }
    }
  }
  return RQ_OK;
}

int rq_bvh_export_wide(rq_ctx* c, const rq_bvh* b, float* boxes, int32_t* links) {
  if (check_ctx(c) || !b || b->magic != BVH_MAGIC) return set_err(c, RQ_E_INVALID, "bad bvh");
  if (!b->B.wide) return set_err(c, RQ_E_UNSUPPORTED, "this BVH has no wide nodes");
  if (b->n < 2) return RQ_OK;
  const size_t ni = b->n - 1;
  std::vector<WideNode> w(ni);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(w.data(), b->B.wide, ni * sizeof(WideNode), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "wide export");
  for (size_t i = 0; i < ni; ++i) {
    const float* q = &w[i].lox.x;  // lox, hix, loy, hiy, loz, hiz (4 slots each)
    for (int k = 0; k < 4 && boxes; ++k) {
      float* o = boxes + 24 * i + 6 * k;
      o[0] = q[k], o[1] = q[8 + k], o[2] = q[16 + k];
      o[3] = q[4 + k], o[4] = q[12 + k], o[5] = q[20 + k];
    }
    if (links) std::memcpy(links + 4 * i, &w[i].link, 16);
  }
  return RQ_OK;
}

// noalloc: the allocation-free entry (rq_closest_hit_noalloc): the device
// scratch array in place of the stream's overflow area and counter slot, the
// grid capped to fit it, no profiling events.
// count_dev / out_index (device arrays, internal callers): the launch traces
// min(n, *count_dev) rays, its slice schedule computed on the device, and
// writes ray r's results at out_index[r]; it takes a counter slot zeroed by a
// kernel, as a captured launch does (its slice count is not known here).
static int query(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, rq_hits* hits,
                 uint8_t* occ, uint32_t* visits, uint32_t* tests, bool any, bool counted,
                 void* stream, bool noalloc = false, const uint32_t* count_dev = nullptr,
                 const uint32_t* out_index = nullptr) {
  if (check_ctx(c) || !b || b->magic != BVH_MAGIC || !rays)
    return set_err(c, RQ_E_INVALID, "bad argument");
  if (n == 0) return RQ_OK;
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz ||
      !rays->tmin || !rays->tmax)
    return set_err(c, RQ_E_INVALID, "null ray array");
  if (any ? !occ : (!hits || !hits->prim)) return set_err(c, RQ_E_INVALID, "null output array");
  if (counted && (!visits || !tests)) return set_err(c, RQ_E_INVALID, "null count array");
  TraceArgs a{};
  a.nodes = b->B.nodes;
  a.tris = b->B.tris;
  a.nprims = b->n;
  a.rays = *rays;
  a.nrays = n;
  if (!any) a.prim = hits->prim, a.t = hits->t, a.u = hits->u, a.v = hits->v;
  a.occ = occ;
  a.visits = visits;
  a.tests = tests;
  a.nrays_dev = count_dev;
  a.out_index = out_index;
  int variant = counted ? 0 : (any ? c->any_variant : c->closest_variant);
  int blk, depth, ts, wide;
  trace_variant_info(variant, &blk, &depth, &ts, &wide);
  if (wide && !b->B.wide) {  // no wide nodes: the canonical binary walk
    variant = 0;
    trace_variant_info(variant, &blk, &depth, &ts, &wide);
  }
  a.wide = b->B.wide;
  // An error left by an earlier call must not read as this launch's (the
  // slot's base below advances only for a launch that was enqueued).
  (void)hipGetLastError();
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing((hipStream_t)stream, &cap);
  if (e != hipSuccess) return hip_err(c, e, "stream capture status");
  const bool captured = cap == hipStreamCaptureStatusActive;
  // a counter slot whose start is not tracked here: zeroed in the stream first
  const bool own_zero = captured || count_dev != nullptr;
  a.gdepth = (uint32_t)(trace_max_stack(variant) - depth);
  if (noalloc) {
    const size_t per_block = (size_t)blk * a.gdepth * 2 * sizeof(int32_t);
    const size_t blocks = (NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES) / per_block;
    if (blocks == 0) return set_err(c, RQ_E_UNSUPPORTED, "variant too large for the scratch");
    a.gstack = reinterpret_cast<int32_t*>(c->scratch + NOALLOC_COUNTER_BYTES);
    a.counter = reinterpret_cast<unsigned long long*>(c->scratch);
    a.ctr_base = 0ull;
    a.err = c->errword;
    a.grid = std::max(1u, std::min({c->trace_grid[variant], (uint32_t)blocks,
                                     (n + (uint32_t)blk - 1) / (uint32_t)blk}));
    a.refill_idle = c->refill_closest;
    const SliceSchedule sched = slice_schedule(
        n, trace_chunk(n, a.grid * (uint32_t)blk / 64u, c->slice_closest));
    a.chunk = sched.chunk, a.chunk_small = sched.small, a.big_slices = sched.big;
    a.sphere = b->scene && b->scene->kind == RQ_PRIM_SPHERE ? 1 : 0;
    const LaunchCtx lc{(hipStream_t)stream, nullptr};
    e = launch_counter_reset(lc, a.counter);
    if (e == hipSuccess) e = launch_trace(lc, a, false, false, variant);
    return e == hipSuccess ? RQ_OK : hip_err(c, e, "trace launch (no-alloc)");
  }
  a.gstack = stream_gstack(c, stream, !captured);
  if (!a.gstack) return captured ? RQ_E_UNSUPPORTED : RQ_E_NOMEM;
  // Eager launches share a slot's monotonic counter and start from its value
  // after the launches enqueued before them (the base). A base baked into a
  // captured graph would be stale at the next replay, so a captured query
  // takes a capture slot of its own, zeroed in the graph before the trace.
  const uint32_t slot = own_zero ? COUNTER_SLOTS + (c->capture_seq++ % CAPTURE_SLOTS)
                                 : next_counter_slot(c);
  a.counter = reinterpret_cast<unsigned long long*>(c->counters) + slot;
  a.ctr_base = own_zero ? 0ull : c->counter_base[slot];
  a.err = c->errword;
  // the launch grid (launch_trace takes the same), its waves and slice size
  const uint32_t full_grid = any ? c->trace_grid_any[variant] : c->trace_grid[variant];
  a.grid = std::max(1u, std::min(full_grid, (n + (uint32_t)blk - 1) / (uint32_t)blk));
  a.refill_idle = any ? c->refill_any : c->refill_closest;
  const uint32_t waves = a.grid * (uint32_t)blk / 64u;
  const SliceSchedule sched =
      slice_schedule(n, trace_chunk(n, full_grid * (uint32_t)blk / 64u,
                                    any ? c->slice_any : c->slice_closest));
  a.chunk = sched.chunk, a.chunk_small = sched.small, a.big_slices = sched.big;
  if (count_dev) a.chunk = any ? c->slice_any : c->slice_closest;  // the largest slice
  a.sphere = b->scene && b->scene->kind == RQ_PRIM_SPHERE ? 1 : 0;
  if (own_zero) {
    e = launch_counter_reset(launch_ctx(c, stream), a.counter);
    if (e != hipSuccess) return hip_err(c, e, "counter reset");
  }
  e = launch_trace(launch_ctx(c, stream), a, any, counted, variant);
  if (e != hipSuccess) return hip_err(c, e, "trace launch");
  // the launch moves its slot's counter by its number of slices, or not at
  // all when the waves' first slices cover it (RayStream)
  if (!own_zero && sched.total > waves) c->counter_base[slot] += sched.total;
  return RQ_OK;
}

extern "C++" {
namespace rq {
int any_hit_indexed(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t capacity,
                    const uint32_t* count_dev, const uint32_t* out_index, uint8_t* occ,
                    void* stream) {
  return query(c, b, rays, capacity, nullptr, occ, nullptr, nullptr, true, false, stream, false,
               count_dev, out_index);
}
}  // namespace rq
}

int rq_closest_hit(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, rq_hits* hits,
                   void* stream) {
  return query(c, b, rays, n, hits, nullptr, nullptr, nullptr, false, false, stream);
}
int rq_closest_hit_noalloc(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                           rq_hits* hits, void* stream) {
  return query(c, b, rays, n, hits, nullptr, nullptr, nullptr, false, false, stream, true);
}
int rq_any_hit(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, uint8_t* occ,
               void* stream) {
  return query(c, b, rays, n, nullptr, occ, nullptr, nullptr, true, false, stream);
}
int rq_closest_hit_counted(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                           rq_hits* hits, uint32_t* visits, uint32_t* tests, void* stream) {
  return query(c, b, rays, n, hits, nullptr, visits, tests, false, true, stream);
}
int rq_any_hit_counted(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, uint8_t* occ,
                       uint32_t* visits, uint32_t* tests, void* stream) {
  return query(c, b, rays, n, nullptr, occ, visits, tests, true, true, stream);
}

int rq_trace_variant_count(void) { return trace_num_variants(); }

int rq_trace_variant_get(int v, rq_trace_variant* out) {
  if (v < 0 || v >= trace_num_variants() || !out) return RQ_E_INVALID;
  trace_variant_info(v, &out->block, &out->lds_depth, &out->tstack, &out->wide,
                     &out->closest_waves, &out->any_waves);
  return RQ_OK;
}

int rq_trace_config_set(rq_ctx* c, int closest_variant, int any_variant, uint32_t refill_idle) {
  if (check_ctx(c)) return RQ_E_INVALID;
  const int nv = trace_num_variants();
  if (closest_variant < 0 || closest_variant >= nv || any_variant < 0 || any_variant >= nv ||
      refill_idle < 1 || refill_idle > 64)
    return set_err(c, RQ_E_INVALID, "bad traversal configuration");
  c->closest_variant = closest_variant;
  c->any_variant = any_variant;
  c->refill_closest = c->refill_any = refill_idle;
  return RQ_OK;
}

int rq_trace_refill_set(rq_ctx* c, uint32_t closest_refill, uint32_t any_refill) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (closest_refill < 1 || closest_refill > 64 || any_refill < 1 || any_refill > 64)
    return set_err(c, RQ_E_INVALID, "refill threshold outside 1..64");
  c->refill_closest = closest_refill;
  c->refill_any = any_refill;
  return RQ_OK;
}

int rq_trace_refill_get(rq_ctx* c, uint32_t* closest_refill, uint32_t* any_refill) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (closest_refill) *closest_refill = c->refill_closest;
  if (any_refill) *any_refill = c->refill_any;
  return RQ_OK;
}

static bool slice_ok(uint32_t s) {
  return s >= 64 && s <= TRACE_RAY_CHUNK_MAX && (s & (s - 1)) == 0;
}

int rq_trace_slice_set(rq_ctx* c, uint32_t closest_slice, uint32_t any_slice) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (!slice_ok(closest_slice) || !slice_ok(any_slice))
    return set_err(c, RQ_E_INVALID, "slice size not a power of two in 64..%u", TRACE_RAY_CHUNK_MAX);
  c->slice_closest = closest_slice;
  c->slice_any = any_slice;
  return RQ_OK;
}

int rq_trace_slice_get(rq_ctx* c, uint32_t* closest_slice, uint32_t* any_slice) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (closest_slice) *closest_slice = c->slice_closest;
  if (any_slice) *any_slice = c->slice_any;
  return RQ_OK;
}

int rq_trace_config_get(rq_ctx* c, int* closest_variant, int* any_variant, uint32_t* refill_idle) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (closest_variant) *closest_variant = c->closest_variant;
  if (any_variant) *any_variant = c->any_variant;
  if (refill_idle) *refill_idle = c->refill_closest;
  return RQ_OK;
}

int rq_autotune_trace(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, int any,
                      int repeats, int* best_variant, uint32_t* best_refill, double* best_ms) {
  if (check_ctx(c) || !b || !rays || n == 0 || repeats < 1) return RQ_E_INVALID;
  // scratch outputs for the timed launches
  void *out0 = nullptr, *out1 = nullptr;
  int rc = ctx_alloc(c, (size_t)n * 4, &out0);
  if (!rc) rc = ctx_alloc(c, (size_t)n * 4, &out1);
  if (rc) {
    ctx_free(c, out0);
    return rc;
  }
  const int saved_c = c->closest_variant, saved_a = c->any_variant;
  const uint32_t saved_rc = c->refill_closest, saved_ra = c->refill_any;
  const uint32_t saved_sc = c->slice_closest, saved_sa = c->slice_any;
  const bool saved_prof = c->prof_on;
  c->prof_on = false;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // mean time of `repeats` launches after an untimed one, in the current configuration
  auto time_it = [&](double* ms_out) {
    rq_hits h{(int32_t*)out0, (float*)out1, nullptr, nullptr};
    for (int k = 0; k <= repeats && !rc; ++k) {
      if (k == 1) (void)hipEventRecord(e0, nullptr);
      rc = any ? rq_any_hit(c, b, rays, n, (uint8_t*)out0, nullptr)
               : rq_closest_hit(c, b, rays, n, &h, nullptr);
    }
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms / repeats;
  };
  double best = 1e30;
  int bv = 0;
  uint32_t br = 16, bs = TRACE_RAY_CHUNK;
  const uint32_t refills[] = {16, 48, 56, 64};
  for (int v = 0; v < trace_num_variants() && !rc; ++v) {
    int blk, depth, ts, wide;
    trace_variant_info(v, &blk, &depth, &ts, &wide);
    if (any && ts) continue;  // entry distances only help closest hit
    if (wide && !b->B.wide) continue;
    for (uint32_t rf : refills) {
      c->closest_variant = v, c->any_variant = v, c->refill_closest = c->refill_any = rf;
      c->slice_closest = c->slice_any = bs;
      double ms = 0;
      time_it(&ms);
      if (!rc && ms < best) best = ms, bv = v, br = rf;
    }
  }
  // then the slice size for the best variant and refill
  const uint32_t slices[] = {256, 512, 2048};
  for (uint32_t sl : slices) {
    if (rc) break;
    c->closest_variant = c->any_variant = bv, c->refill_closest = c->refill_any = br;
    c->slice_closest = c->slice_any = sl;
    double ms = 0;
    time_it(&ms);
    if (!rc && ms < best) best = ms, bs = sl;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx_free(c, out0);
  ctx_free(c, out1);
  c->prof_on = saved_prof;
  c->closest_variant = saved_c, c->any_variant = saved_a;
  c->refill_closest = saved_rc, c->refill_any = saved_ra;
  c->slice_closest = saved_sc, c->slice_any = saved_sa;
  if (rc) return rc;
  if (any)
    c->any_variant = bv, c->refill_any = br, c->slice_any = bs;
  else
    c->closest_variant = bv, c->refill_closest = br, c->slice_closest = bs;
  if (best_variant) *best_variant = bv;
  if (best_refill) *best_refill = br;
  if (best_ms) *best_ms = best;
  return RQ_OK;
}

int rq_device_alloc(rq_ctx* c, size_t bytes, void** p) {
  if (check_ctx(c) || !p) return RQ_E_INVALID;
  hipSetDevice(c->device);
  return ctx_alloc(c, bytes, p);
}
int rq_device_free(rq_ctx* c, void* p) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (p && !c->sizes.count(p)) return set_err(c, RQ_E_INVALID, "pointer not owned by context");
  ctx_free(c, p);
  return RQ_OK;
}
int rq_memcpy_h2d(rq_ctx* c, void* dst, const void* src, size_t bytes, void* stream) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "memcpy h2d");
}
int rq_memcpy_d2h(rq_ctx* c, void* dst, const void* src, size_t bytes, void* stream) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "memcpy d2h");
}
int rq_memcpy_d2d(rq_ctx* c, void* dst, const void* src, size_t bytes, void* stream) {
  if (check_ctx(c) || (bytes && (!dst || !src))) return RQ_E_INVALID;
  if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
    return set_err(c, RQ_E_INVALID, "device copy needs 16-byte aligned pointers and size");
  hipError_t e = launch_copy16(launch_ctx(c, stream), src, dst, bytes);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "device copy");
}
int rq_memset_d(rq_ctx* c, void* dst, int value, size_t bytes, void* stream) {
  hipError_t e = hipMemsetAsync(dst, value, bytes, (hipStream_t)stream);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "memset");
}
int rq_sync(rq_ctx* c, void* stream) {
  if (check_ctx(c)) return RQ_E_INVALID;
  // NULL: every stream of the device (the library's streams do not
  // synchronise with the null stream)
  (void)hipSetDevice(c->device);
  hipError_t e = stream ? hipStreamSynchronize((hipStream_t)stream) : hipDeviceSynchronize();
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "stream sync");
}
int rq_stream_create(rq_ctx* c, void** stream) {
  if (check_ctx(c) || !stream) return RQ_E_INVALID;
  hipSetDevice(c->device);
  hipStream_t s;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_err(c, e, "stream create");
  *stream = (void*)s;
  return RQ_OK;
}
int rq_stream_forget(rq_ctx* c, void* stream) {
  if (check_ctx(c) || !stream) return RQ_E_INVALID;
  // its overflow area is freed once its work is done (a later stream may get
  // the same handle)
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  for (size_t i = 0; i < c->gstack_streams.size(); ++i)
    if (c->gstack_streams[i].first == stream) {
      ctx_free(c, c->gstack_streams[i].second);
      c->gstack_streams.erase(c->gstack_streams.begin() + (long)i);
      break;
    }
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "stream forget");
}

int rq_stream_order(rq_ctx* c, void* waiter, void* signaler) {
  if (check_ctx(c)) return RQ_E_INVALID;
  if (waiter == signaler) return RQ_OK;  // one stream runs in order
  constexpr size_t kEvents = 16;
  if (c->order_events.empty()) {
    (void)hipSetDevice(c->device);
    for (size_t i = 0; i < kEvents; ++i) {
      hipEvent_t e = nullptr;
      const hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
      if (r != hipSuccess) return hip_err(c, r, "order event");
      c->order_events.push_back(e);
    }
  }
  hipEvent_t ev = c->order_events[c->order_seq++ % kEvents];
  hipError_t e = hipEventRecord(ev, (hipStream_t)signaler);
  if (e == hipSuccess) e = hipStreamWaitEvent((hipStream_t)waiter, ev, 0);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "stream order");
}

int rq_stream_destroy(rq_ctx* c, void* stream) {
  if (check_ctx(c)) return RQ_E_INVALID;
  const int rc = stream ? rq_stream_forget(c, stream) : RQ_OK;
  if (rc) return rc;
  const hipError_t e = hipStreamDestroy((hipStream_t)stream);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "stream destroy");
}

int rq_device_errors(rq_ctx* c, uint32_t* bits) {
  if (check_ctx(c) || !bits) return RQ_E_INVALID;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(bits, c->errword, 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && *bits) e = hipMemset(c->errword, 0, 4);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "device error word");
}

int rq_profile_enable(rq_ctx* c, int on) {
  if (check_ctx(c)) return RQ_E_INVALID;
  c->prof_on = on != 0;
  return RQ_OK;
}
int rq_profile_only(rq_ctx* c, const char* names) {
  if (check_ctx(c)) return RQ_E_INVALID;
  c->prof->only.clear();
  if (!names) return RQ_OK;
  std::string cur;
  for (const char* q = names;; ++q) {
    if (*q == ',' || *q == '\0') {
      if (!cur.empty()) c->prof->only.insert(cur);
      cur.clear();
      if (*q == '\0') break;
    } else {
      cur.push_back(*q);
    }
  }
  return RQ_OK;
}
int rq_profile_reset(rq_ctx* c) {
  if (check_ctx(c)) return RQ_E_INVALID;
  c->prof->resolve();
  c->prof->totals.clear();
  return RQ_OK;
}
int rq_profile_read(rq_ctx* c, char names[][48], double* total_ms, uint64_t* launches, int cap) {
  if (check_ctx(c)) return RQ_E_INVALID;
  c->prof->resolve();
  int k = 0;
  for (auto& kv : c->prof->totals) {
    if (k < cap) {
      std::snprintf(names[k], 48, "%s", kv.first.c_str());
      total_ms[k] = kv.second.first;
      launches[k] = kv.second.second;
    }
    ++k;
  }
  return k;
}

int rq_release(rq_ctx* c, void* h) {
  if (check_ctx(c) || !h) return RQ_E_INVALID;
  uint32_t magic;
  std::memcpy(&magic, h, 4);
  if (magic == SCENE_MAGIC) {
    auto* s = (rq_scene*)h;
    s->magic = 0;
    if (s->mat) ctx_free(c, s->mat);
    ctx_free(c, s->base);
    delete s;
    return RQ_OK;
  }
  if (magic == BVH_MAGIC) {
    auto* b = (rq_bvh*)h;
    if (!b->owned) return set_err(c, RQ_E_INVALID, "BVH belongs to a batch");
    b->magic = 0;
    ctx_free(c, b->base);
    delete b;
    return RQ_OK;
  }
  return set_err(c, RQ_E_INVALID, "unknown handle");
}

}  // extern "C"
