// Internal definitions shared by the HIP kernels and the host side of librq.
// Not part of the public ABI (include/rq/*.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "rq/rq.h"

namespace rq {

// ---------------------------------------------------------------------------
// Device data layout
// ---------------------------------------------------------------------------
// Internal node: the boxes of BOTH children plus the two child links, so one
// visit is one 64-byte record (four 16-byte loads). Child link >= 0 is an
// internal node index, < 0 is leaf ~k (k = sorted position).
struct alignas(16) NodeRec {
  float4 lo0;  // lo of child 0, .w = left link (int bits)
  float4 hi0;  // hi of child 0, .w = right link (int bits)
  float4 lo1;  // lo of child 1, .w unused
  float4 hi1;  // hi of child 1, .w unused
};
static_assert(sizeof(NodeRec) == 64, "node record is 64 bytes");

// Leaf triangle in sorted order: v0, e1 = v1 - v0, e2 = v2 - v0 (48 bytes).
struct alignas(16) TriRec {
  float4 v0;  // .w = primitive id (int bits)
  float4 e1;
  float4 e2;
};
static_assert(sizeof(TriRec) == 48, "triangle record is 48 bytes");

// 4-wide node: the binary LBVH collapsed two levels at a time. The wide node
// of a binary internal node i (kept at index i; every internal node has one,
// the walk reaches the root and the nodes at even depth) holds the boxes of
// i's grandchildren, or of a child that is a leaf, as four child slots in SoA
// order, one 128-byte record per visit. Links follow the binary convention
// (>= 0 wide node = binary index, < 0 leaf ~k); an unused slot has link 0
// (the root is never a child) and an empty box (lo +inf, hi -inf).
struct alignas(128) WideNode {
  float4 lox, hix, loy, hiy, loz, hiz;
  int4 link;
  int4 pad;
};
static_assert(sizeof(WideNode) == 128, "wide node record is 128 bytes");

// Radix sort tiling (onesweep): WAVES*64 threads, ITEMS keys per thread.
// The standalone build sorts 16 keys per thread; the batch kernel, whose
// register budget is set by traversal, sorts 8 (SORT_ITEMS_LEAN). Look-back
// status words are allocated for the smaller tile, so either fits.
constexpr int SORT_WAVES = 4;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_ITEMS_LEAN = 8;
constexpr int SORT_BLOCK = SORT_WAVES * 64;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;
constexpr int SORT_TILE_LEAN = SORT_BLOCK * SORT_ITEMS_LEAN;
constexpr int SORT_PASSES = 4;  // 30-bit codes, 8-bit digits

// Refit: a block owns REFIT_CHUNK consecutive leaves (the standalone kernel
// one leaf per thread, the batch kernel four); the top pass runs in one
// work-group.
constexpr int REFIT_CHUNK = 1024;
// The smallest refit chunk (the batch kernel's 256-thread blocks): the
// chunk-union buffers hold one union per chunk of this size.
constexpr int REFIT_CHUNK_MIN = 256;

// Lookback status word: [31:30] flag, [29:0] count.
constexpr uint32_t LB_AGG = 1u << 30;
constexpr uint32_t LB_INC = 2u << 30;
constexpr uint32_t LB_MASK = (1u << 30) - 1;
constexpr uint32_t MAX_PRIMS = (1u << 30) - 1;

// Traversal: the LBVH of 30-bit codes is at most 63 levels deep (prefix
// lengths 2..63 along a path), so 64 stack entries per lane always suffice;
// kernel variants keep 12..32 of them in LDS and the rest in a global
// overflow area ([depth][thread] pairs of {node, entry distance}).
constexpr int TRACE_MAX_STACK = 64;
// Wide traversal keeps up to 3 entries per wide ancestor, and a path of at
// most 64 binary levels has at most 32 wide ones.
constexpr int TRACE_MAX_STACK_WIDE = 96;
constexpr int TRACE_MIN_LDS_DEPTH = 12;
constexpr int TRACE_MAX_VARIANTS = 24;
// Variants a new context uses: the 4-wide walk, 128 threads and 24 LDS entries
// for closest hit, 256 threads and 16 entries for any hit (the fastest or
// within a few per cent on camera, shadow, bounce and uniform random rays).
constexpr int TRACE_DEFAULT_CLOSEST = 9;
constexpr int TRACE_DEFAULT_ANY = 14;
constexpr int COUNTER_SLOTS = 64;
// Counter slots of queries enqueued on a stream under graph capture: each
// captured query zeroes its own slot in the graph before its trace, so every
// replay starts from 0 (round robin: at most CAPTURE_SLOTS captured queries
// per context may run at once).
constexpr int CAPTURE_SLOTS = 64;
constexpr int MAX_TRACE_STREAMS = 16;  // streams besides the default one with traversal launches
constexpr uint32_t TRACE_RAY_CHUNK = 1024;      // default largest wave slice of the ray stream
constexpr uint32_t TRACE_RAY_CHUNK_MAX = 4096;  // largest that may be configured

// Rays per wave slice for a launch of n rays over `waves` persistent waves:
// `largest` (a power of two, 64..4096) when every wave gets at least two full
// slices, halved down to one ray per lane (64) for smaller launches, so that
// a small batch still spreads over as many waves as it can fill (each slice
// after a wave's first costs one device atomic on the launch's counter).
__host__ __device__ inline uint32_t trace_chunk(uint32_t n, uint32_t waves,
                                                uint32_t largest = TRACE_RAY_CHUNK) {
  uint32_t c = largest;
  while (c > 64 && (uint64_t)n < 2ull * c * waves) c >>= 1;
  return c;
}
// Slice schedule of a standalone launch: slices of `chunk` rays over the
// first three quarters of the stream, then slices of chunk / 4 (at least 64)
// so that the waves finish together; slice s starts at slice_begin(s).
struct SliceSchedule {
  uint32_t chunk, small, big, total;  // sizes, number of large slices, number of slices
};
__host__ __device__ inline SliceSchedule slice_schedule(uint32_t n, uint32_t chunk) {
  SliceSchedule q;
  q.chunk = chunk;
  q.small = chunk / 4 < 64 ? 64 : chunk / 4;
  q.big = (uint32_t)((3ull * n / 4) / chunk);
  const uint64_t rest = n - (uint64_t)q.big * chunk;
  q.total = q.big + (uint32_t)((rest + q.small - 1) / q.small);
  return q;
}

// Build workspace (device), all sized for n primitives.
struct BuildBuffers {
  uint32_t* codes = nullptr;      // n, primitive order
  uint32_t* keys_a = nullptr;     // n
  uint32_t* keys_b = nullptr;     // n
  uint32_t* vals_a = nullptr;     // n  (final sorted_idx)
  uint32_t* vals_b = nullptr;     // n
  uint32_t* status = nullptr;     // SORT_PASSES * ntiles * 256 (tiles of SORT_TILE_LEAN)
  uint32_t* misc = nullptr;       // see MiscLayout
  int32_t* parent = nullptr;      // 2n-1: (node << 1) | side
  uint32_t* flags = nullptr;      // n-1 words, zeroed by Karras: tree-height and treelet scratch
  uint32_t* slot_h = nullptr;     // 2(n-1) words of treelet scratch (heights, counts)
  uint2* range = nullptr;         // n-1 leaf range [first, last] of each internal node
  uint32_t* frontier = nullptr;   // n words of treelet scratch (treelet roots)
  NodeRec* nodes = nullptr;       // n-1
  TriRec* tris = nullptr;         // n
  WideNode* wide = nullptr;       // n-1 (see WideNode)
  // Chunk-union refit. Boxes are 6 floats (lo xyz, hi xyz).
  float* pre = nullptr;           // 6n: union of the leaves from the chunk start to k
  float* suf = nullptr;           // 6n: union of the leaves from k to the chunk end
  float* cbox = nullptr;          // 6 per refit chunk (up to n / REFIT_CHUNK_MIN): its leaves' union
  uint32_t ntiles = 0;            // tiles of SORT_TILE_LEAN keys
};

// misc[] layout (uint32 words)
constexpr int MISC_HIST = 0;                          // SORT_PASSES * 256
constexpr int MISC_TILECTR = MISC_HIST + SORT_PASSES * 256;  // SORT_PASSES
constexpr int MISC_BOUNDS = MISC_TILECTR + SORT_PASSES;      // 6 ordered-uint (lo xyz, hi xyz)
constexpr int MISC_ROOT = MISC_BOUNDS + 6;                   // 6 floats root box
constexpr int MISC_HEIGHT = MISC_ROOT + 6;                   // 1 (HEIGHT_UNKNOWN: not computed)
constexpr int MISC_ERR = MISC_HEIGHT + 1;                    // 1: build error bits
constexpr int MISC_WORDS = MISC_ERR + 1;
constexpr uint32_t HEIGHT_UNKNOWN = 0xFFFFFFFFu;

// Build error bits (never expected; they turn a logic error into a reported
// failure instead of an out-of-range access or an endless wait).
constexpr uint32_t ERR_LOOKBACK_SPIN = 1u;
constexpr uint32_t ERR_SCATTER_RANGE = 2u;
constexpr uint32_t ERR_TOPOLOGY = 4u;
constexpr uint32_t ERR_REFIT_DEPTH = 8u;
constexpr uint32_t ERR_WIDEN_DEPTH = 16u;
constexpr uint32_t LOOKBACK_SPIN_LIMIT = 1u << 24;

// ---------------------------------------------------------------------------
// Host-side launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
struct Profiler;
struct LaunchCtx {
  hipStream_t stream;
  Profiler* prof;  // may be null
};

// vr: sphere radii (spheres centred at vx/vy/vz), or null for triangles.
hipError_t launch_build(const LaunchCtx& lc, const float* vx, const float* vy, const float* vz,
                        const uint32_t* idx, uint32_t n, int morton_mode, BuildBuffers& B,
                        const float* vr = nullptr);
// Writes B.wide from the binary nodes of a built LBVH (after a blob load; the
// standalone build writes it during its refit).
hipError_t launch_widen(const LaunchCtx& lc, BuildBuffers& B, uint32_t n);
// Computes the tree height into misc (the chunk-union refit leaves it open).
hipError_t launch_height(const LaunchCtx& lc, BuildBuffers& B, uint32_t n);
// Treelet restructuring of a built LBVH in place (`rounds` rounds of 7-leaf
// treelets); `height` is the tree's height before. Leaves the wide nodes and
// the height to launch_widen / launch_height.
hipError_t launch_treelets(const LaunchCtx& lc, BuildBuffers& B, uint32_t n, uint32_t height,
                           int rounds);

struct TraceArgs {
  const NodeRec* nodes;
  const WideNode* wide;  // wide variants only
  const TriRec* tris;
  uint32_t nprims;
  rq_rays rays;
  uint32_t nrays;
  int32_t* prim;
  float *t, *u, *v;
  uint8_t* occ;
  uint32_t* visits;
  uint32_t* tests;
  int32_t* gstack;
  uint32_t gdepth;      // overflow entries per thread
  unsigned long long* counter;  // ray-stream counter of this launch's slot (never reset)
  uint64_t ctr_base;    // its value when this launch starts
  uint32_t* err;        // set to 1 if a ray needed a deeper stack than provided
  uint32_t grid;        // blocks (persistent grid of the chosen variant)
  uint32_t refill_idle; // idle lanes in a wave before it refills from its slice
  uint32_t chunk;       // rays per slice taken from the global counter (the large slices)
  uint32_t chunk_small; // rays per slice after the first big_slices (standalone kernels)
  uint32_t big_slices;  // slices of `chunk` rays before the small ones
  int32_t sphere;       // leaves are spheres (centre, r, r*r) instead of triangles
  // Set when the ray count is known on the device only: the launch traces rays
  // [0, min(nrays, *nrays_dev)) and computes its slice schedule from that
  // count (`chunk` is then the largest slice; the counter starts at zero).
  const uint32_t* nrays_dev;
  // Set when results go elsewhere than to the ray's own index: ray r's
  // outputs are written at out_index[r].
  const uint32_t* out_index;
};
hipError_t launch_trace(const LaunchCtx& lc, const TraceArgs& a, bool any, bool counted,
                        int variant);
int trace_num_variants();
void trace_variant_info(int v, int* block, int* depth, int* tstack, int* wide = nullptr,
                        int* closest_waves = nullptr, int* any_waves = nullptr);
int trace_max_stack(int variant);  // total stack entries per lane (LDS + overflow)
uint32_t trace_grid_blocks(int device, int variant, bool any);
hipError_t launch_copy16(const LaunchCtx& lc, const void* src, void* dst, uint64_t bytes);
// Zeroes one ray-stream counter with a one-thread kernel (a memset node would
// not be safe to replay in a captured graph).
hipError_t launch_counter_reset(const LaunchCtx& lc, unsigned long long* counter);
// Scratch of the allocation-free closest-hit entry: NOALLOC_SCRATCH_BYTES per
// device (a static device array), its first NOALLOC_COUNTER_BYTES the
// counter, the rest the overflow area; the launch grid is capped so that the
// overflow area of every thread fits.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = 256;
// The device address of the scratch on the current device.
hipError_t noalloc_scratch_address(void** ptr);

// Profiler hooks (rq_api.cpp)
void prof_begin(Profiler* p, hipStream_t s, const char* name, hipEvent_t* ev);
void prof_end(Profiler* p, hipStream_t s, const char* name, hipEvent_t ev);

struct ProfScope {
  Profiler* p;
  hipStream_t s;
  const char* name;
  hipEvent_t ev = nullptr;
  ProfScope(const LaunchCtx& lc, const char* n) : p(lc.prof), s(lc.stream), name(n) {
    if (p) prof_begin(p, s, name, &ev);
  }
  ~ProfScope() {
    if (p) prof_end(p, s, name, ev);
  }
};

}  // namespace rq
