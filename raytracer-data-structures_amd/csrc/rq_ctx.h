// Host-side handle types behind the opaque ABI pointers.
#pragma once
#include <string>
#include <utility>
#include <vector>
#include <unordered_map>

#include "rq_internal.h"

struct rq_ctx {
  int device = 0;
  size_t budget = 0;
  size_t used = 0;
  std::string err;
  std::unordered_map<void*, size_t> sizes;  // live device allocations
  rq::Profiler* prof = nullptr;
  bool prof_on = false;
  // persistent grid per kernel variant and query kind (closest, any)
  uint32_t trace_grid[rq::TRACE_MAX_VARIANTS] = {}, trace_grid_any[rq::TRACE_MAX_VARIANTS] = {};
  // chosen traversal variants (the 4-wide walks until tuned or set)
  int closest_variant = rq::TRACE_DEFAULT_CLOSEST, any_variant = rq::TRACE_DEFAULT_ANY;
  // idle lanes before a wave refills (64: only once every lane is idle), per query kind
  uint32_t refill_closest = 64, refill_any = 64;
  // largest wave slice of the ray stream, per query kind (trace_chunk)
  uint32_t slice_closest = rq::TRACE_RAY_CHUNK, slice_any = rq::TRACE_RAY_CHUNK;
  // COUNTER_SLOTS 64-bit ray-stream counters (never reset), their values
  // after the launches enqueued so far, and one word of traversal error bits
  uint32_t* counters = nullptr;
  uint64_t counter_base[rq::COUNTER_SLOTS] = {};
  uint32_t* errword = nullptr;   // traversal error bits (stack deeper than provided)
  uint32_t counter_seq = 0;
  uint32_t capture_seq = 0;      // next capture slot (after the COUNTER_SLOTS eager ones)
  // Traversal stack overflow ({node, distance} pairs), indexed by global
  // thread id: launches that may run at once need separate areas, so each
  // stream gets its own (the default stream's is allocated with the context).
  int32_t* gstack = nullptr;
  size_t gstack_bytes = 0;
  std::vector<std::pair<void*, int32_t*>> gstack_streams;
  // events of rq_stream_order, reused round robin (a wait enqueued on a
  // stream refers to the record that preceded it)
  std::vector<hipEvent_t> order_events;
  uint32_t order_seq = 0;
  // staging of the host-pointer queries (host_query.cpp): a pinned host
  // buffer and a device buffer (charged to the budget) of stage_bytes each
  void* stage_host = nullptr;
  void* stage_dev = nullptr;
  size_t stage_bytes = 0;
  // device address of the allocation-free entry's scratch on this device
  // (hipGetSymbolAddress at rq_create)
  unsigned char* scratch = nullptr;
};

constexpr uint32_t SCENE_MAGIC = 0x53435131u;  // "SCQ1"
constexpr uint32_t BVH_MAGIC = 0x42565131u;    // "BVQ1"
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct rq_scene {
  uint32_t magic = 0;  // first member: handle type tag
  uint32_t nv = 0, ntri = 0;  // spheres: nv = ntri = number of spheres
  int kind = 0;               // RQ_PRIM_TRIANGLE / RQ_PRIM_SPHERE
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;  // vertices, or sphere centres
  uint32_t* idx = nullptr;                            // triangles only
  float* r = nullptr;                                 // sphere radii
  float* rgb = nullptr;                               // optional surface colour per primitive
  uint8_t* mat = nullptr;  // optional material per primitive (own allocation)
  void* base = nullptr;
  size_t bytes = 0;
};

struct rq_bvh {
  uint32_t magic = 0;  // first member: handle type tag
  const rq_scene* scene = nullptr;
  int mode = 0;
  uint32_t n = 0;
  rq::BuildBuffers B;
  void* base = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: part of a batch, released with it
};

namespace rq {
int set_err(rq_ctx* c, int code, const char* fmt, ...);
int hip_err(rq_ctx* c, hipError_t e, const char* what);
int ctx_alloc(rq_ctx* c, size_t bytes, void** p);
void ctx_free(rq_ctx* c, void* p);
LaunchCtx launch_ctx(rq_ctx* c, void* stream);
// The next ray-stream counter slot (round robin over COUNTER_SLOTS). A launch
// passes counter_base[slot] to its kernel and then adds the number of slices
// its waves take from the counter (RayStream), so the slot's next launch
// starts from the counter's value after this one.
uint32_t next_counter_slot(rq_ctx* c);
// The stack overflow area of launches on `stream`, allocated on a stream's
// first traversal launch (not under graph capture: `may_alloc` false then);
// null with the context's error set when the budget or the limit of
// MAX_TRACE_STREAMS streams is exceeded, or when allocation is not allowed.
// At the limit, the area of a stream with no work left is taken over.
int32_t* stream_gstack(rq_ctx* c, void* stream, bool may_alloc = true);
// Build buffers for n primitives: the binary and 4-wide node arrays, the
// triangle records and the chunk-union refit's unions (every BVH: standalone,
// blob-loaded, batch meshes and server slots).
int bvh_alloc(rq_ctx* c, rq_bvh* b, uint32_t n);
// Makes the tree height in b's misc words known (blocking; the chunk-union
// refit leaves it to this call).
int ensure_height(rq_ctx* c, const rq_bvh* b);
// Frees the pinned staging buffer of the host-pointer queries (rq_destroy).
void stage_release(rq_ctx* c);
// Any hit over min(capacity, *count_dev) rays whose count is known on the
// device only, ray r's flag written at occ[out_index[r]] (the renderer's
// compacted second-hit shadow rays).
int any_hit_indexed(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t capacity,
                    const uint32_t* count_dev, const uint32_t* out_index, uint8_t* occ,
                    void* stream);
}  // namespace rq
