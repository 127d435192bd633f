// Serving loop (include/rq/rq_serve.h): resident mesh slots, an arena for
// request buffers, FIFO admission against the arena / in-flight limit / slot
// pool, continuous batching (every step is one trace-only persistent launch
// over ray tiles of every ready request), and result copies on their own
// stream. One host thread runs the loop and makes every HIP call of the
// server; submit / poll / wait only touch the request table under a mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_internal.h"
#include "rq/rq_serve.h"
#include "rq_ctx.h"
#include "rq_render_internal.h"

using namespace rq;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation: offsets aligned to 256 B,
// free neighbours coalesced on release.
class Arena {
 public:
  void init(size_t bytes) {
    free_.clear();
    if (bytes) free_[0] = bytes;
    cap_ = bytes;
    used = peak = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used += bytes;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align256(bytes);
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return cap_; }
  size_t used = 0, peak = 0;

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0;
};

struct Slot {
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  uint32_t* idx = nullptr;
  void* base = nullptr;
  rq_bvh bvh;
  uint32_t vcap = 0, tcap = 0, n = 0;
  bool loaded = false;
  bool resident = false;  // loaded by rq_server_load_mesh: never recycled
  uint32_t pins = 0;      // queued or admitted requests that trace this slot
  uint64_t last_use = 0;
};

struct Req {
  uint64_t ticket = 0;
  rq_request r{};
  int32_t state = RQ_REQ_QUEUED;
  int32_t err = 0;
  uint32_t slot = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  uint32_t steps = 0;
  size_t off = 0, bytes = 0;  // arena range
  float* ray[8] = {};
  int32_t* prim = nullptr;
  float* t = nullptr;
  uint8_t* occ = nullptr;
  hipEvent_t prep = nullptr, out = nullptr;
  bool pinned = false;        // holds a pin on `slot`
  bool resident = false;      // a mesh load of rq_server_load_mesh (a slot of its own)
  std::atomic<bool> cancel{false};  // rq_server_cancel: finish as RQ_E_CANCELED
  bool ready = false;         // rays (and a new mesh) are on the device
  uint32_t assigned = 0;      // rays handed to steps
  uint32_t pending = 0;       // chunks in steps not yet retired
  bool out_enqueued = false;
  // every output array is page-locked host memory: each chunk's results are
  // copied when its step is launched (behind the step's end event); a
  // pageable array would make such a copy block the loop until the step ends,
  // so those requests get one copy once their last chunk is traced
  bool stream_out = false;
};

struct Chunk {
  Req* q;
  uint32_t first, count;
};

struct Step {
  hipEvent_t start = nullptr, end = nullptr;
  std::vector<Chunk> chunks;
  bool active = false;
  uint64_t seq = 0;
};

constexpr int kStepBufs = 2;           // steps in flight (double-buffered tables)
constexpr uint32_t kMinStepTile = 256;  // smallest ray tile of a step (one ray per thread)

// Whether p points to page-locked host memory (hipHostMalloc / rq_host_alloc).
bool page_locked(const void* p) {
  if (!p) return true;
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // a pageable pointer is not an error of a later call
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

size_t request_bytes(const rq_request& r) {
  const size_t n = r.nrays;
  size_t b = 0;
  if (n) b += 8 * align256(4 * n);
  if (r.query & RQ_Q_CLOSEST) b += align256(4 * n) * (r.t ? 2 : 1);
  if (r.query & RQ_Q_ANY) b += align256(n);
  return b;
}

}  // namespace

struct rq_server {
  rq_server_config cfg{};
  int device = 0;
  rq_ctx* ctx = nullptr;
  std::vector<Slot> slots;
  Arena arena;
  char* arena_base = nullptr;
  hipStream_t s_prep = nullptr, s_trace = nullptr, s_out = nullptr;
  BatchMeshDev* dev_meshes[kStepBufs] = {};
  BatchWork* dev_items[kStepBufs] = {};
  BatchMeshDev* host_meshes[kStepBufs] = {};
  BatchWork* host_items[kStepBufs] = {};
  uint32_t* err_host[kStepBufs] = {};
  size_t max_items = 0;
  uint32_t* ctl = nullptr;
  int32_t* gstack = nullptr;
  uint32_t gdepth = 0, grid = 0;
  // trace kernel of the steps (serve_variant_info): the 4-wide walk, LDS
  // stack depth 16, 6 waves per SIMD (against the binary walk, serving 64
  // meshes of 10k-2M triangles at 2000 requests/s: p50 1.40 -> 1.15 ms,
  // saturation 2.90 -> 3.16 Grays/s)
  int variant = 7;
  Step steps[kStepBufs];
  uint64_t step_seq = 0;

  std::mutex mu;
  std::condition_variable cv_loop, cv_done;
  std::deque<Req*> incoming;                              // submitted, not yet seen by the loop
  std::deque<uint64_t> finished;                          // tickets not yet taken by take_done
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;  // every known request (owned)
  uint64_t next_ticket = 1;
  uint32_t resident_slots = 0;  // slots taken (or claimed by queued loads) by load_mesh
  bool stop = false;
  std::thread th;
  rq_server_stats stats{};
  std::string err;

  // loop-thread state
  std::deque<Req*> waiting;
  std::vector<Req*> admitted;
  uint64_t use_clock = 0;
};

namespace {

int srv_err(rq_server* s, int code, const std::string& m) {
  std::lock_guard<std::mutex> g(s->mu);
  s->err = m;
  return code;
}

// Publishes a request's new state (loop thread).
void publish(rq_server* s, Req* q, int32_t state) {
  std::lock_guard<std::mutex> g(s->mu);
  q->state = state;
  if (state == RQ_REQ_DONE || state < 0) {
    if (state < 0)
      ++s->stats.failed;
    else
      ++s->stats.completed;
    const bool had_slot = q->pinned;
    if (q->pinned) {
      --s->slots[q->slot].pins;
      s->slots[q->slot].last_use = ++s->use_clock;
      q->pinned = false;
    }
    if (q->resident && state < 0) {
      --s->resident_slots;  // a load that did not happen gives its claim back
      if (had_slot) s->slots[q->slot].resident = false, s->slots[q->slot].loaded = false;
    }
    s->finished.push_back(q->ticket);
    s->cv_done.notify_all();
  }
}

// A slot for a new mesh: an empty one, else the least recently used idle
// slot that holds a per-request upload. A resident mesh (rq_server_load_mesh)
// keeps its slot for the server's life, so a slot id a caller holds always
// names the mesh it loaded.
int pick_slot(rq_server* s) {
  std::lock_guard<std::mutex> g(s->mu);
  int best = -1;
  for (size_t i = 0; i < s->slots.size(); ++i) {
    const Slot& sl = s->slots[i];
    if (sl.pins || sl.resident) continue;
    if (!sl.loaded) return (int)i;
    if (best < 0 || sl.last_use < s->slots[best].last_use) best = (int)i;
  }
  return best;
}

// Camera basis on the host: f = normalize(look_at - eye), r = normalize(f x up),
// u = r x f; a w x ceil(n / w) image.
PinholeParams make_pinhole(const rq_pinhole& c, uint32_t n, uint64_t seed) {
  float f[3] = {c.look_at[0] - c.eye[0], c.look_at[1] - c.eye[1], c.look_at[2] - c.eye[2]};
  auto nrm = [](float* v) {
    const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0f) v[0] /= l, v[1] /= l, v[2] /= l;
  };
  nrm(f);
  float rt[3] = {f[1] * c.up[2] - f[2] * c.up[1], f[2] * c.up[0] - f[0] * c.up[2],
                 f[0] * c.up[1] - f[1] * c.up[0]};
  nrm(rt);
  const float up[3] = {rt[1] * f[2] - rt[2] * f[1], rt[2] * f[0] - rt[0] * f[2],
                       rt[0] * f[1] - rt[1] * f[0]};
  PinholeParams pp{};
  std::memcpy(pp.eye, c.eye, sizeof pp.eye);
  std::memcpy(pp.f, f, sizeof f);
  std::memcpy(pp.r, rt, sizeof rt);
  std::memcpy(pp.u, up, sizeof up);
  const uint32_t w = std::max<uint32_t>(1, c.width);
  const uint32_t h = std::max<uint32_t>(1, (uint32_t)(((uint64_t)n + w - 1) / w));
  pp.w = w;
  pp.n = n;
  pp.tan_half = std::tan(0.5f * c.fov_deg * 3.14159265358979f / 180.0f);
  pp.aspect = (float)w / (float)h;
  pp.inv_w = 1.0f / (float)w;
  pp.inv_h = 1.0f / (float)h;
  pp.seed = seed;
  return pp;
}

BuildView slot_view(const Slot& sl) {
  return BuildView{sl.vx, sl.vy, sl.vz, sl.idx, sl.n, RQ_MORTON_SCENE_AABB, sl.bvh.B};
}

// Admission in submission order. A request that does not fit the arena or
// the in-flight limit stops admission (first come, first served); a new-mesh
// request that finds every slot pinned is passed over, so that the requests
// holding the pins can run and release them.
void admit(rq_server* s) {
  for (auto wi = s->waiting.begin();
       wi != s->waiting.end() && s->admitted.size() < s->cfg.max_inflight;) {
    Req* q = *wi;
    const rq_request& r = q->r;
    if (q->cancel.load()) {
      wi = s->waiting.erase(wi);
      q->t_done = now_s();
      publish(s, q, RQ_E_CANCELED);
      continue;
    }
    size_t off = 0;
    if (q->bytes && !s->arena.alloc(q->bytes, &off)) break;
    if (r.new_mesh) {
      const int sl = pick_slot(s);
      if (sl < 0) {
        if (q->bytes) s->arena.release(off, q->bytes);
        ++wi;
        continue;
      }
      std::lock_guard<std::mutex> g(s->mu);
      q->slot = (uint32_t)sl;
      q->pinned = true;
      ++s->slots[sl].pins;
    }
    wi = s->waiting.erase(wi);
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stats.arena_peak = std::max(s->stats.arena_peak, s->arena.peak);
    }
    q->off = off;
    char* p = s->arena_base + off;
    const size_t n = r.nrays;
    if (n) {
      for (int k = 0; k < 8; ++k) q->ray[k] = (float*)p, p += align256(4 * n);
      if (r.query & RQ_Q_CLOSEST) {
        q->prim = (int32_t*)p, p += align256(4 * n);
        if (r.t) q->t = (float*)p, p += align256(4 * n);
      }
      if (r.query & RQ_Q_ANY) q->occ = (uint8_t*)p;
    }
    const LaunchCtx lc{s->s_prep, nullptr};
    hipError_t e = hipSuccess;
    if (r.new_mesh) {
      Slot& sl = s->slots[q->slot];
      const rq_mesh& m = *r.new_mesh;
      {
        std::lock_guard<std::mutex> g(s->mu);
        sl.loaded = true;
        sl.resident = q->resident;
      }
      sl.n = m.ntri;
      e = hipMemcpyAsync(sl.vx, m.vx, 4ull * m.nv, hipMemcpyHostToDevice, s->s_prep);
      if (e == hipSuccess) e = hipMemcpyAsync(sl.vy, m.vy, 4ull * m.nv, hipMemcpyHostToDevice, s->s_prep);
      if (e == hipSuccess) e = hipMemcpyAsync(sl.vz, m.vz, 4ull * m.nv, hipMemcpyHostToDevice, s->s_prep);
      if (e == hipSuccess)
        e = hipMemcpyAsync(sl.idx, m.idx, 12ull * m.ntri, hipMemcpyHostToDevice, s->s_prep);
      sl.bvh.n = m.ntri;
      if (e == hipSuccess)
        e = launch_build(lc, sl.vx, sl.vy, sl.vz, sl.idx, m.ntri, RQ_MORTON_SCENE_AABB, sl.bvh.B);
      ++s->stats.builds;
    }
    if (e == hipSuccess && n) {
      const RaySoA R{q->ray[0], q->ray[1], q->ray[2], q->ray[3],
                     q->ray[4], q->ray[5], q->ray[6], q->ray[7]};
      if (r.source == RQ_RAYS_HOST) {
        const float* src[8] = {r.host_rays.ox, r.host_rays.oy, r.host_rays.oz,   r.host_rays.dx,
                               r.host_rays.dy, r.host_rays.dz, r.host_rays.tmin, r.host_rays.tmax};
        for (int k = 0; k < 8 && e == hipSuccess; ++k)
          e = hipMemcpyAsync(q->ray[k], src[k], 4 * n, hipMemcpyHostToDevice, s->s_prep);
      } else if (r.source == RQ_RAYS_UNIFORM) {
        e = launch_raygen_uniform(lc, (uint32_t)n, r.seed, r.lo, r.hi, R);
      } else {
        e = launch_raygen_pinhole(lc, make_pinhole(r.pinhole, (uint32_t)n, r.seed), R);
      }
    }
    if (e == hipSuccess) e = hipEventRecord(q->prep, s->s_prep);
    if (e != hipSuccess) q->err = RQ_E_HIP;  // completes as failed once nothing uses it
    q->t_admit = now_s();
    s->admitted.push_back(q);
    publish(s, q, RQ_REQ_ADMITTED);
  }
}

void poll_ready(rq_server* s) {
  for (Req* q : s->admitted) {
    if (q->ready) continue;
    const hipError_t e = hipEventQuery(q->prep);
    if (e == hipErrorNotReady) continue;
    q->ready = true;
    if (e != hipSuccess) q->err = RQ_E_HIP;
  }
}

// Forms and launches one step if a table buffer is free and work is ready.
bool launch_step(rq_server* s) {
  Step* st = nullptr;
  int buf = -1;
  for (int b = 0; b < kStepBufs; ++b)
    if (!s->steps[b].active) {
      st = &s->steps[b];
      buf = b;
      break;
    }
  if (!st) return false;
  st->chunks.clear();
  uint64_t total = 0;
  for (Req* q : s->admitted) {
    if (!q->ready || q->err || q->cancel.load() || q->assigned >= q->r.nrays) continue;
    if (st->chunks.size() >= s->cfg.max_inflight || total >= s->cfg.step_rays) break;
    const uint32_t c = std::min<uint32_t>(q->r.nrays - q->assigned, s->cfg.chunk_rays);
    st->chunks.push_back(Chunk{q, q->assigned, c});
    total += c;
  }
  if (st->chunks.empty()) return false;
  // Ray tile of this step: halved from the configured one (down to 256 rays,
  // one per thread) until the step has work items for 4 blocks per grid
  // slot, so that a small step still spreads over the whole device.
  auto count_items = [&](uint32_t tl) {
    size_t k = 0;
    for (const Chunk& c : st->chunks)
      k += (size_t)((c.count + tl - 1) / tl) * ((c.q->prim ? 1 : 0) + (c.q->occ ? 1 : 0));
    return k;
  };
  uint32_t tile = s->cfg.ray_tile;
  while (tile > kMinStepTile && count_items(tile) < 4ull * s->grid) tile >>= 1;
  // mesh table and work items (ray tiles dealt to XCD segments, LPT by rays)
  BatchMeshDev* md = s->host_meshes[buf];
  const uint32_t ne = (uint32_t)st->chunks.size();
  std::vector<uint64_t> work(ne);
  for (uint32_t i = 0; i < ne; ++i) {
    const Chunk& c = st->chunks[i];
    Req* q = c.q;
    BatchMeshDev& d = md[i];
    d = BatchMeshDev{};
    d.v = slot_view(s->slots[q->slot]);
    d.rays = rq_rays{q->ray[0] + c.first, q->ray[1] + c.first, q->ray[2] + c.first,
                     q->ray[3] + c.first, q->ray[4] + c.first, q->ray[5] + c.first,
                     q->ray[6] + c.first, q->ray[7] + c.first};
    d.nrays = c.count;
    d.prim = q->prim ? q->prim + c.first : nullptr;
    d.t = q->t ? q->t + c.first : nullptr;
    d.occ = q->occ ? q->occ + c.first : nullptr;
    work[i] = (uint64_t)c.count * ((d.prim ? 1 : 0) + (d.occ ? 1 : 0));
  }
  std::vector<uint32_t> order(ne), seg(ne);
  for (uint32_t i = 0; i < ne; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
  uint64_t load[BATCH_XCDS] = {};
  for (uint32_t i : order) {
    const int qd = (int)(std::min_element(load, load + BATCH_XCDS) - load);
    seg[i] = (uint32_t)qd;
    load[qd] += work[i];
  }
  BatchArgs a{};
  BatchWork* it = s->host_items[buf];
  uint32_t ni = 0;
  for (int qd = 0; qd < BATCH_XCDS; ++qd) {
    a.trace_seg[qd] = ni;
    for (uint32_t i = 0; i < ne; ++i) {
      if (seg[i] != (uint32_t)qd) continue;
      const uint32_t nt = (md[i].nrays + tile - 1) / tile;
      if (md[i].prim)
        for (uint32_t t = 0; t < nt; ++t) it[ni++] = BatchWork{i, t};
      if (md[i].occ)
        for (uint32_t t = 0; t < nt; ++t) it[ni++] = BatchWork{i, t | BATCH_ANY_BIT};
    }
  }
  a.trace_seg[BATCH_XCDS] = ni;
  a.meshes = s->dev_meshes[buf];
  a.items = s->dev_items[buf];
  a.ctl = s->ctl;
  a.gstack = s->gstack;
  a.gdepth = s->gdepth;
  a.ray_tile = tile;
  a.refill_idle = 16;
  hipError_t e = hipMemcpyAsync(s->dev_meshes[buf], md, ne * sizeof(BatchMeshDev),
                                hipMemcpyHostToDevice, s->s_trace);
  if (e == hipSuccess && ni)
    e = hipMemcpyAsync(s->dev_items[buf], it, ni * sizeof(BatchWork), hipMemcpyHostToDevice,
                       s->s_trace);
  if (e == hipSuccess) e = hipEventRecord(st->start, s->s_trace);
  if (e == hipSuccess) e = launch_serve_trace(LaunchCtx{s->s_trace, nullptr}, a, s->variant, s->grid);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->err_host[buf], s->ctl + BATCH_CTL_BARRIER_ERR, 2 * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, s->s_trace);
  if (e == hipSuccess) e = hipEventRecord(st->end, s->s_trace);
  // Each chunk's results go to the caller's (page-locked) host arrays as soon
  // as the step ends: copies on the output stream behind the step's end
  // event, so they overlap the later steps, and a request's last copy is its
  // last chunk's, queued when that chunk is launched (not when the loop sees
  // the step end).
  if (e == hipSuccess) e = hipStreamWaitEvent(s->s_out, st->end, 0);
  for (const Chunk& c : st->chunks) {
    Req* q = c.q;
    if (!q->stream_out) continue;
    if (e == hipSuccess && q->prim)
      e = hipMemcpyAsync(q->r.prim + c.first, q->prim + c.first, 4ull * c.count,
                         hipMemcpyDeviceToHost, s->s_out);
    if (e == hipSuccess && q->t)
      e = hipMemcpyAsync(q->r.t + c.first, q->t + c.first, 4ull * c.count, hipMemcpyDeviceToHost,
                         s->s_out);
    if (e == hipSuccess && q->occ)
      e = hipMemcpyAsync(q->r.occluded + c.first, q->occ + c.first, c.count,
                         hipMemcpyDeviceToHost, s->s_out);
    if (e == hipSuccess) e = hipEventRecord(q->out, s->s_out);
    q->out_enqueued = true;
  }
  for (const Chunk& c : st->chunks) {
    c.q->assigned += c.count;
    ++c.q->pending;
    ++c.q->steps;
  }
  st->active = true;
  st->seq = ++s->step_seq;
  {
    std::lock_guard<std::mutex> g(s->mu);
    ++s->stats.steps;
    s->stats.max_step_requests = std::max(s->stats.max_step_requests, ne);
  }
  if (e != hipSuccess) {
    for (const Chunk& c : st->chunks) c.q->err = RQ_E_HIP;
  }
  return true;
}

// Retires finished steps in launch order. The result copies of page-locked
// requests were queued at launch; a pageable request gets one copy of all its
// results once its last chunk has retired.
void enqueue_out(rq_server* s, Req* q, hipEvent_t after) {
  const size_t n = q->r.nrays;
  hipError_t e = hipStreamWaitEvent(s->s_out, after, 0);
  if (e == hipSuccess && q->prim)
    e = hipMemcpyAsync(q->r.prim, q->prim, 4 * n, hipMemcpyDeviceToHost, s->s_out);
  if (e == hipSuccess && q->t)
    e = hipMemcpyAsync(q->r.t, q->t, 4 * n, hipMemcpyDeviceToHost, s->s_out);
  if (e == hipSuccess && q->occ)
    e = hipMemcpyAsync(q->r.occluded, q->occ, n, hipMemcpyDeviceToHost, s->s_out);
  if (e == hipSuccess) e = hipEventRecord(q->out, s->s_out);
  if (e != hipSuccess) q->err = RQ_E_HIP;
  q->out_enqueued = true;
}

void retire_steps(rq_server* s) {
  for (;;) {
    Step* st = nullptr;
    for (auto& x : s->steps)
      if (x.active && (!st || x.seq < st->seq)) st = &x;
    if (!st) return;
    const hipError_t e = hipEventQuery(st->end);
    if (e == hipErrorNotReady) return;
    const int buf = (int)(st - s->steps);
    const bool bad = e != hipSuccess || s->err_host[buf][0] || s->err_host[buf][1];
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, st->start, st->end);
    uint64_t rays = 0;
    for (const Chunk& c : st->chunks) {
      Req* q = c.q;
      --q->pending;
      if (bad) q->err = e != hipSuccess ? RQ_E_HIP : RQ_E_DEVICE;
      rays += (uint64_t)c.count * ((q->prim ? 1 : 0) + (q->occ ? 1 : 0));
      if (!q->stream_out && q->pending == 0 && q->assigned >= q->r.nrays && !q->out_enqueued &&
          !q->err && !q->cancel.load())
        enqueue_out(s, q, st->end);
    }
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stats.busy_s += ms * 1e-3;
      s->stats.rays += rays;
    }
    st->active = false;
  }
}

// A request is finished when no step still holds one of its chunks and its
// preparation and result copies are over; only then are its arena range and
// slot pin released (the device may still be writing to them before that).
void complete(rq_server* s) {
  for (size_t i = 0; i < s->admitted.size();) {
    Req* q = s->admitted[i];
    bool done = false;
    if (q->ready && q->pending == 0) {
      // the result copies queued so far are over (they read the arena range)
      bool copied = true;
      if (q->out_enqueued) {
        const hipError_t e = hipEventQuery(q->out);
        if (e == hipErrorNotReady)
          copied = false;
        else if (e != hipSuccess)
          q->err = RQ_E_HIP;
      }
      if (!copied) {
      } else if (q->err || q->r.nrays == 0) {
        done = true;  // failed, or a mesh load without rays
      } else if (q->cancel.load()) {
        q->err = RQ_E_CANCELED, done = true;  // no step holds it, no copy is running
      } else if (q->assigned >= q->r.nrays && q->out_enqueued) {
        done = true;  // every chunk traced and copied
      }
    }
    if (!done) {
      ++i;
      continue;
    }
    if (q->bytes) s->arena.release(q->off, q->bytes), q->bytes = 0;
    q->t_done = now_s();
    publish(s, q, q->err ? q->err : RQ_REQ_DONE);
    s->admitted.erase(s->admitted.begin() + (long)i);
  }
}

void loop(rq_server* s) {
  (void)hipSetDevice(s->device);
  auto last_progress = std::chrono::steady_clock::now();
  for (;;) {
    bool stopping;
    {
      std::unique_lock<std::mutex> g(s->mu);
      if (s->incoming.empty() && s->waiting.empty() && s->admitted.empty()) {
        bool idle_steps = true;
        for (auto& x : s->steps) idle_steps &= !x.active;
        if (idle_steps) {
          if (s->stop) break;
          s->cv_loop.wait_for(g, std::chrono::milliseconds(50));
        }
      }
      while (!s->incoming.empty()) {
        s->waiting.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
      stopping = s->stop;
    }
    if (stopping) {
      // refuse what has not been admitted; finish what has
      while (!s->waiting.empty()) {
        Req* q = s->waiting.front();
        s->waiting.pop_front();
        q->t_done = now_s();
        publish(s, q, RQ_E_INVALID);
      }
    }
    const size_t before = s->admitted.size() + s->waiting.size();
    const uint64_t steps_before = s->step_seq;
    admit(s);
    poll_ready(s);
    retire_steps(s);
    while (launch_step(s)) {
    }
    complete(s);
    const bool progress = s->admitted.size() + s->waiting.size() != before ||
                          s->step_seq != steps_before;
    const auto now = std::chrono::steady_clock::now();
    if (progress) {
      last_progress = now;
    } else if (now - last_progress > std::chrono::milliseconds(2)) {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    } else {
      std::this_thread::yield();
    }
  }
}

void free_server(rq_server* s) {
  if (s->ctx) {
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    for (auto& st : s->steps) {
      if (st.start) (void)hipEventDestroy(st.start);
      if (st.end) (void)hipEventDestroy(st.end);
    }
    for (auto& kv : s->reqs) {
      if (kv.second->prep) (void)hipEventDestroy(kv.second->prep);
      if (kv.second->out) (void)hipEventDestroy(kv.second->out);
    }
    for (int b = 0; b < kStepBufs; ++b) {
      if (s->host_meshes[b]) (void)hipHostFree(s->host_meshes[b]);
      if (s->host_items[b]) (void)hipHostFree(s->host_items[b]);
      if (s->err_host[b]) (void)hipHostFree(s->err_host[b]);
    }
    for (hipStream_t st : {s->s_prep, s->s_trace, s->s_out})
      if (st) (void)hipStreamDestroy(st);
    rq_destroy(s->ctx);  // frees every device allocation of the server
  }
  delete s;
}

}  // namespace

extern "C" {

void rq_server_defaults(rq_server_config* c) {
  if (!c) return;
  c->hbm_budget = (size_t)200 * 1000 * 1000 * 1000;
  c->mesh_slots = 64;
  c->slot_tris = 1u << 21;
  c->arena_bytes = (size_t)8 << 30;
  c->max_inflight = 64;
  c->step_rays = 1u << 22;
  c->chunk_rays = 1u << 20;
  c->ray_tile = 4096;
}

double rq_server_clock(void) { return now_s(); }

int rq_raygen_pinhole(rq_ctx* c, const rq_pinhole* cam, uint32_t n, uint64_t seed, float* ox,
                      float* oy, float* oz, float* dx, float* dy, float* dz, float* tmin,
                      float* tmax, void* stream) {
  if (!c || !cam || (n && (!ox || !oy || !oz || !dx || !dy || !dz || !tmin || !tmax)))
    return RQ_E_INVALID;
  (void)hipSetDevice(c->device);
  const hipError_t e = launch_raygen_pinhole(launch_ctx(c, stream), make_pinhole(*cam, n, seed),
                                             RaySoA{ox, oy, oz, dx, dy, dz, tmin, tmax});
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "raygen_pinhole");
}

void* rq_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void rq_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int rq_server_create(int device, const rq_server_config* cfg_in, rq_server** out) {
  if (!out) return RQ_E_INVALID;
  *out = nullptr;
  rq_server_config cfg;
  if (cfg_in)
    cfg = *cfg_in;
  else
    rq_server_defaults(&cfg);
  if (!cfg.mesh_slots || !cfg.slot_tris || cfg.slot_tris > MAX_PRIMS || !cfg.max_inflight ||
      !cfg.step_rays || !cfg.chunk_rays || cfg.ray_tile < 256)
    return RQ_E_INVALID;
  auto* s = new rq_server();
  s->cfg = cfg;
  s->device = device;
  int rc = rq_create(&s->ctx, device, cfg.hbm_budget);
  if (rc) {
    s->ctx = nullptr;
    delete s;
    return rc;
  }
  rq_ctx* c = s->ctx;
  // mesh slots: vertices (up to 3 per triangle), indices, LBVH build buffers
  s->slots.resize(cfg.mesh_slots);
  for (auto& sl : s->slots) {
    if (rc) break;
    sl.tcap = cfg.slot_tris;
    sl.vcap = 3 * cfg.slot_tris;
    const size_t vb = align256(4ull * sl.vcap), ib = align256(12ull * sl.tcap);
    rc = ctx_alloc(c, 3 * vb + ib, &sl.base);
    if (rc) break;
    char* p = (char*)sl.base;
    sl.vx = (float*)p, sl.vy = (float*)(p + vb), sl.vz = (float*)(p + 2 * vb);
    sl.idx = (uint32_t*)(p + 3 * vb);
    sl.bvh.magic = BVH_MAGIC;
    sl.bvh.owned = false;
    sl.bvh.mode = RQ_MORTON_SCENE_AABB;
    rc = bvh_alloc(c, &sl.bvh, sl.tcap);
  }
  if (!rc) rc = ctx_alloc(c, std::max<size_t>(cfg.arena_bytes, 256), (void**)&s->arena_base);
  s->arena.init(cfg.arena_bytes);
  s->max_items =
      2 * ((size_t)(cfg.step_rays + cfg.chunk_rays) / kMinStepTile + cfg.max_inflight) + 16;
  for (int b = 0; b < kStepBufs && !rc; ++b) {
    rc = ctx_alloc(c, cfg.max_inflight * sizeof(BatchMeshDev), (void**)&s->dev_meshes[b]);
    if (!rc) rc = ctx_alloc(c, s->max_items * sizeof(BatchWork), (void**)&s->dev_items[b]);
    if (!rc && (hipHostMalloc((void**)&s->host_meshes[b], cfg.max_inflight * sizeof(BatchMeshDev),
                              hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&s->host_items[b], s->max_items * sizeof(BatchWork),
                              hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&s->err_host[b], 2 * sizeof(uint32_t),
                              hipHostMallocDefault) != hipSuccess))
      rc = set_err(c, RQ_E_NOMEM, "pinned staging");
    if (!rc) {
      s->err_host[b][0] = s->err_host[b][1] = 0;
      if (hipEventCreate(&s->steps[b].start) != hipSuccess ||
          hipEventCreate(&s->steps[b].end) != hipSuccess)
        rc = RQ_E_HIP;
    }
  }
  if (!rc) rc = ctx_alloc(c, BATCH_CTL_ALLOC_WORDS * sizeof(uint32_t), (void**)&s->ctl);
  if (!rc && hipMemset(s->ctl, 0, BATCH_CTL_ALLOC_WORDS * sizeof(uint32_t)) != hipSuccess) rc = RQ_E_HIP;
  if (!rc) {
    int depth = 0, ts = 0, wide = 0;
    serve_variant_info(s->variant, &depth, &ts, &wide);
    s->grid = serve_grid_blocks(device, s->variant);
    s->gdepth = (wide ? TRACE_MAX_STACK_WIDE : TRACE_MAX_STACK) - depth;
    rc = ctx_alloc(c, (size_t)s->grid * BUILD_BLOCK * s->gdepth * 2 * sizeof(int32_t),
                   (void**)&s->gstack);
  }
  for (hipStream_t* st : {&s->s_prep, &s->s_trace, &s->s_out})
    if (!rc && hipStreamCreateWithFlags(st, hipStreamNonBlocking) != hipSuccess) rc = RQ_E_HIP;
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = RQ_E_HIP;
  if (rc) {
    free_server(s);
    return rc;
  }
  s->th = std::thread(loop, s);
  *out = s;
  return RQ_OK;
}

int rq_server_destroy(rq_server* s) {
  if (!s) return RQ_E_INVALID;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_loop.notify_all();
  if (s->th.joinable()) s->th.join();
  free_server(s);
  return RQ_OK;
}

}  // extern "C"

namespace {

// resident: a mesh load of rq_server_load_mesh, which takes a slot for the
// server's life; other new-mesh requests upload into recycled slots.
int submit(rq_server* s, const rq_request* r, uint64_t* ticket, bool resident) {
  if (!s || !r || !ticket) return RQ_E_INVALID;
  const bool closest = r->query & RQ_Q_CLOSEST, any = r->query & RQ_Q_ANY;
  if ((r->query & ~(RQ_Q_CLOSEST | RQ_Q_ANY)) || (r->nrays && !closest && !any) ||
      (!r->nrays && !r->new_mesh))
    return srv_err(s, RQ_E_INVALID, "request: no rays, no query or an unknown query bit");
  if ((closest && !r->prim) || (any && !r->occluded))
    return srv_err(s, RQ_E_INVALID, "request: missing output array");
  if (r->nrays && r->source != RQ_RAYS_HOST && r->source != RQ_RAYS_PINHOLE &&
      r->source != RQ_RAYS_UNIFORM)
    return srv_err(s, RQ_E_INVALID, "request: unknown ray source");
  if (r->nrays && r->source == RQ_RAYS_HOST &&
      (!r->host_rays.ox || !r->host_rays.oy || !r->host_rays.oz || !r->host_rays.dx ||
       !r->host_rays.dy || !r->host_rays.dz || !r->host_rays.tmin || !r->host_rays.tmax))
    return srv_err(s, RQ_E_INVALID, "request: null host ray array");
  if (r->new_mesh) {
    const rq_mesh& m = *r->new_mesh;
    if (!m.vx || !m.vy || !m.vz || !m.idx || !m.ntri)
      return srv_err(s, RQ_E_INVALID, "request: empty new mesh");
    if (m.ntri > s->cfg.slot_tris || m.nv > 3ull * s->cfg.slot_tris)
      return srv_err(s, RQ_E_NOMEM, "request: mesh larger than a slot");
    for (uint64_t k = 0; k < 3ull * m.ntri; ++k)
      if (m.idx[k] >= m.nv) return srv_err(s, RQ_E_INVALID, "request: index out of range");
  }
  auto q = std::make_unique<Req>();
  q->r = *r;
  q->bytes = request_bytes(*r);
  if (q->bytes > s->arena.capacity())
    return srv_err(s, RQ_E_NOMEM, "request: buffers larger than the admission arena");
  q->t_submit = now_s();
  {
    // events are created on the caller's thread: hipEventCreate is thread-safe
    (void)hipSetDevice(s->device);
    if (hipEventCreateWithFlags(&q->prep, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&q->out, hipEventDisableTiming) != hipSuccess)
      return srv_err(s, RQ_E_HIP, "request events");
    q->stream_out = page_locked(r->prim) && page_locked(r->t) && page_locked(r->occluded);
  }
  std::lock_guard<std::mutex> g(s->mu);
  auto refuse = [&](int code, const char* why) {
    (void)hipEventDestroy(q->prep);
    (void)hipEventDestroy(q->out);
    s->err = why;
    return code;
  };
  if (s->stop) return refuse(RQ_E_INVALID, "server stopping");
  if (!r->new_mesh) {
    if (r->mesh >= s->slots.size() || !s->slots[r->mesh].loaded || !s->slots[r->mesh].resident)
      return refuse(RQ_E_INVALID,
                    "request: mesh slot holds no resident mesh (an uploaded mesh serves only "
                    "the request that brought it)");
    q->slot = r->mesh;
    q->pinned = true;
    ++s->slots[r->mesh].pins;
  } else if (s->resident_slots >= s->slots.size()) {
    return refuse(RQ_E_NOMEM, resident ? "every mesh slot holds a resident mesh"
                                       : "no mesh slot for uploads: every slot holds a "
                                         "resident mesh");
  } else if (resident) {
    ++s->resident_slots;
    q->resident = true;
  }
  q->ticket = s->next_ticket++;
  *ticket = q->ticket;
  s->incoming.push_back(q.get());
  s->reqs.emplace(q->ticket, std::move(q));
  ++s->stats.submitted;
  s->cv_loop.notify_one();
  return RQ_OK;
}

}  // namespace

extern "C" {

int rq_server_submit(rq_server* s, const rq_request* r, uint64_t* ticket) {
  return submit(s, r, ticket, false);
}

int rq_server_cancel(rq_server* s, uint64_t ticket) {
  if (!s) return RQ_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return RQ_E_INVALID;
  Req* q = it->second.get();
  if (q->state == RQ_REQ_DONE || q->state < 0) return RQ_E_INVALID;  // already finished
  if (q->resident) return RQ_E_INVALID;  // a mesh load of rq_server_load_mesh
  q->cancel.store(true);
  s->cv_loop.notify_one();
  return RQ_OK;
}

static void fill_status(const Req* q, rq_request_status* st) {
  st->state = q->state;
  st->mesh = q->slot;
  st->steps = q->steps;
  st->t_submit = q->t_submit;
  st->t_admit = q->t_admit;
  st->t_done = q->t_done;
}

int rq_server_poll(rq_server* s, uint64_t ticket, rq_request_status* st) {
  if (!s || !st) return RQ_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return RQ_E_INVALID;
  fill_status(it->second.get(), st);
  return RQ_OK;
}

int rq_server_wait(rq_server* s, uint64_t ticket, rq_request_status* st) {
  if (!s || !st) return RQ_E_INVALID;
  std::unique_lock<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return RQ_E_INVALID;
  Req* q = it->second.get();
  s->cv_done.wait(g, [&] { return q->state == RQ_REQ_DONE || q->state < 0; });
  fill_status(q, st);
  return q->state < 0 ? q->state : RQ_OK;
}

int rq_server_take_done(rq_server* s, uint64_t* tickets, rq_request_status* st, uint32_t cap,
                        uint32_t* count) {
  if (!s || !count || (cap && (!tickets || !st))) return RQ_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  uint32_t k = 0;
  while (k < cap && !s->finished.empty()) {
    const uint64_t t = s->finished.front();
    s->finished.pop_front();
    auto it = s->reqs.find(t);
    if (it == s->reqs.end()) continue;  // forgotten already
    tickets[k] = t;
    fill_status(it->second.get(), &st[k]);
    ++k;
  }
  *count = k;
  return RQ_OK;
}

int rq_server_forget(rq_server* s, uint64_t ticket) {
  if (!s) return RQ_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return RQ_E_INVALID;
  Req* q = it->second.get();
  if (q->state != RQ_REQ_DONE && q->state >= 0) return RQ_E_INVALID;  // still in flight
  (void)hipEventDestroy(q->prep);
  (void)hipEventDestroy(q->out);
  q->prep = q->out = nullptr;
  s->reqs.erase(it);
  return RQ_OK;
}

int rq_server_load_mesh(rq_server* s, const rq_mesh* m, uint32_t* slot) {
  if (!s || !m || !slot) return RQ_E_INVALID;
  rq_request r{};
  r.new_mesh = m;
  uint64_t t = 0;
  int rc = submit(s, &r, &t, true);
  if (rc) return rc;
  rq_request_status st{};
  rc = rq_server_wait(s, t, &st);
  *slot = st.mesh;
  rq_server_forget(s, t);
  return rc;
}

int rq_server_stats_get(rq_server* s, rq_server_stats* st) {
  if (!s || !st) return RQ_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  *st = s->stats;
  uint32_t res = 0;
  for (const Slot& sl : s->slots) res += sl.loaded ? 1 : 0;
  st->resident = res;
  return RQ_OK;
}

const char* rq_server_last_error(rq_server* s) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> g(s->mu);
  return s->err.c_str();
}

}  // extern "C"
