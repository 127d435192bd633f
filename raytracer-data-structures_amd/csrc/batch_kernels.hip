// Persistent batch kernel: builds the LBVH of many meshes and traces their
// ray batches in one launch. Blocks pull work items (mesh, tile) from a device
// work queue, one queue segment per phase (init, bounds, Morton, four sort
// passes, Karras, refit, refit top, trace) holding the items of every mesh;
// a grid-wide barrier separates phases, so a mesh's phase k + 1 reads what
// any block wrote in phase k. Within a phase the queue balances uneven meshes:
// a block takes the next item as soon as it finishes one. Items of a sort pass
// are queued mesh by mesh in tile order, so a tile's look-back only waits on
// tiles already taken by running blocks.
#include <algorithm>

#include "batch_internal.h"
#include "build_device.h"
#include "trace_device.h"

namespace rq {
namespace {

// s_getreg operand of HW_REG_XCC_ID (register 20, bits [3:0]): the XCD id.
constexpr int XCC_ID_HWREG = 20 | (0 << 6) | ((4 - 1) << 11);

// About 4M polls (a few seconds): far beyond any phase of the member blocks.
constexpr uint32_t BARRIER_SPIN_LIMIT = 1u << 22;
// Polls the first barrier waits for the rest of the grid to start before it
// closes the membership with the blocks that have.
constexpr uint32_t JOIN_SPIN_LIMIT = 1u << 14;

// Membership: the barriers count only the blocks that have started, so that
// they never wait for a block the device has not placed (kernels on other
// streams can hold CUs when the launch begins). A block joins by taking a
// ticket from the start word at entry; a block that finds the membership
// closed leaves at once. The first block to reach the first barrier waits
// (bounded) until the whole grid has started, then closes the membership
// with an atomic OR, whose returned count is the number of members: the
// members hold tickets 0 .. members - 1. Work is pulled from queues, so any
// number of members completes every phase: a grid that could not be placed
// whole runs on fewer blocks, not wrongly. Returns the block's ticket, or
// ~0u for a block that takes no part.
__device__ inline uint32_t join_grid(uint32_t* ctl, uint32_t* s_flag) {
  if (threadIdx.x == 0) *s_flag = atomicAdd(ctl + BATCH_CTL_START, 1u);
  __syncthreads();
  const uint32_t t = *s_flag;
  __syncthreads();
  return (t & BATCH_CLOSED) ? ~0u : t;
}

__device__ inline uint32_t* bar_group(uint32_t* ctl, uint32_t g) {
  return ctl + BATCH_BAR_BASE + (size_t)g * BATCH_BAR_STRIDE;
}
__device__ inline uint32_t* bar_release(uint32_t* ctl, uint32_t g) {
  return ctl + BATCH_BAR_BASE + (size_t)(BATCH_BAR_GROUPS + g) * BATCH_BAR_STRIDE;
}

// Phase stamps (BATCH_CTL_CLOCK): the latest (or, for the start, earliest)
// wall-clock reading over the member blocks, one atomic per block.
__device__ inline unsigned long long* clock_stamp(uint32_t* ctl, uint32_t k) {
  return reinterpret_cast<unsigned long long*>(ctl + BATCH_CTL_CLOCK) + k;
}

// Grid barrier after phase `phases_done` - 1. The block's waves meet at a
// work-group barrier, which leaves their stores in the XCD's L2; thread 0's
// agent-scope release then writes that L2 back (one write-back per block
// rather than one per wave), and it counts the block on its group's counter
// (group = ticket mod BATCH_BAR_GROUPS). The group's last arrival counts the
// group on the top counter, and the top's last arrival publishes the phase
// on every group's release word; each block waits on its own group's word
// and acquires. The counters are cumulative over the launch, so the last
// arrival of phase p is the one that brings a counter to p times its
// members. One word shared by every block, for the arrivals and for the
// polls, had made each barrier cost ~0.15-0.3 ms on a full grid. The
// members are running, so every wait ends; it is bounded all the same: a
// wait that gives up sets the sticky error word, and once it is set no later
// barrier of the launch waits at all.
__device__ inline void grid_barrier(uint32_t* ctl, uint32_t phases_done, uint32_t ticket) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    uint32_t* const err = ctl + BATCH_CTL_BARRIER_ERR;
    const uint32_t g = ticket % (uint32_t)BATCH_BAR_GROUPS;
    const uint32_t old = atomicAdd(bar_group(ctl, g), 1u);
    uint32_t members = ld_agent(ctl + BATCH_CTL_MEMBERS);
    if (members == 0u) {  // the first barrier of the launch: close the membership
      uint32_t spins = 0;
      for (;;) {
        const uint32_t st = ld_agent(ctl + BATCH_CTL_START);
        if ((st & BATCH_CLOSED) || st >= gridDim.x || ++spins > JOIN_SPIN_LIMIT) break;
        __builtin_amdgcn_s_sleep(2);
      }
      const uint32_t o = atomicOr(ctl + BATCH_CTL_START, BATCH_CLOSED);
      if (!(o & BATCH_CLOSED)) atomicExch(ctl + BATCH_CTL_MEMBERS, o);
      spins = 0;
      while ((members = ld_agent(ctl + BATCH_CTL_MEMBERS)) == 0u) {
        if (++spins > BARRIER_SPIN_LIMIT) {
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    if (members != 0u) {
      constexpr uint32_t G = BATCH_BAR_GROUPS;
      const uint32_t groups = min(members, G), in_group = (members - g + G - 1u) / G;
      if (old + 1u == in_group * phases_done) {  // the group's last arrival
        __threadfence();
        if (atomicAdd(ctl + BATCH_CTL_BARRIER, 1u) + 1u == groups * phases_done) {
          __threadfence();  // the grid's last arrival: release every group
          for (uint32_t k = 0; k < G; ++k) atomicMax(bar_release(ctl, k), phases_done);
        }
      }
      uint32_t spins = 0;
      while (ld_agent(bar_release(ctl, g)) < phases_done) {
        if (++spins > BARRIER_SPIN_LIMIT || ((spins & 1023u) == 0u && ld_agent(err) != 0u)) {
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();
    atomicMax(clock_stamp(ctl, phases_done), (unsigned long long)wall_clock64());
  }
  __syncthreads();
}

// Loads a block-uniform descriptor into scalar registers: every dword goes
// through readfirstlane, so the struct is provably uniform and is never copied
// to per-lane scratch.
template <typename T>
__device__ inline T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized descriptor");
  constexpr int W = sizeof(T) / 4;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
  uint32_t w[W];
#pragma unroll
  for (int i = 0; i < W; ++i) w[i] = __builtin_amdgcn_readfirstlane(src[i]);
  T t;
  __builtin_memcpy(&t, w, sizeof(T));
  return t;
}

// One trace item (a tile of one mesh's rays): a ray counter in LDS hands out
// chunks of the tile as lanes go idle. WAVE_ITEM: one wave takes the item on
// its own (its own counter, no block barrier at either end); otherwise the
// block's waves share it and meet at a barrier when it is done.
template <int DEPTH, bool TSTACK, bool WIDE, bool WAVE_ITEM>
__device__ inline void trace_item(int32_t* gstack, uint32_t gdepth, uint32_t* err,
                                  uint32_t ray_tile, uint32_t refill_idle, const BatchMeshDev& m,
                                  uint32_t tile, uint32_t* lds, uint32_t* s_ctr) {
  const bool any = (tile & BATCH_ANY_BIT) != 0;
  const uint32_t r0 = (tile & ~BATCH_ANY_BIT) * ray_tile;
  const uint32_t nr = min(ray_tile, m.nrays - r0);
  TraceArgs t{};
  t.nodes = m.v.B.nodes;
  t.wide = m.v.B.wide;
  t.tris = m.v.B.tris;
  t.nprims = m.v.n;
  t.rays = rq_rays{m.rays.ox + r0, m.rays.oy + r0, m.rays.oz + r0, m.rays.dx + r0,
                   m.rays.dy + r0, m.rays.dz + r0, m.rays.tmin + r0, m.rays.tmax + r0};
  t.nrays = nr;
  t.gstack = gstack;
  t.gdepth = gdepth;
  t.err = err;
  t.refill_idle = refill_idle;
  t.chunk = 64;
  t.sphere = 0;  // batch meshes are triangle scenes
  if (WAVE_ITEM) {
    // lane 0 resets the counter and is the lane that takes from it (trace_loop)
    if ((threadIdx.x & 63) == 0) *s_ctr = 0u;
  } else {
    if (threadIdx.x == 0) *s_ctr = 0u;
    __syncthreads();
  }
  int32_t* s_node = reinterpret_cast<int32_t*>(lds);
  float* s_tn = reinterpret_cast<float*>(lds + DEPTH * BUILD_BLOCK);
  const uint32_t gtid = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  const uint32_t gstride = gridDim.x * BUILD_BLOCK;
  RayStream stream;
  stream.ctr32 = s_ctr;
  if (any) {
    t.occ = m.occ + r0;
    trace_loop<true, false, BUILD_BLOCK, DEPTH, false, false, WIDE>(t, stream, s_node, s_tn,
                                                                    gtid, gstride);
  } else {
    t.prim = m.prim + r0;
    t.t = m.t ? m.t + r0 : nullptr;
    t.u = m.u ? m.u + r0 : nullptr;
    t.v = m.v_ ? m.v_ + r0 : nullptr;
    trace_loop<false, false, BUILD_BLOCK, DEPTH, TSTACK, false, WIDE>(t, stream, s_node, s_tn,
                                                                      gtid, gstride);
  }
  if (!WAVE_ITEM) __syncthreads();
}

// Slices of one wave's walk over one mesh and query kind (the batch run): the
// tile of the item the walk starts with, then the tiles of the items the wave
// takes next from the XCD segments while they are of the same mesh and kind.
// A mesh's items sit in a row in its segment (closest-hit tiles, then any-hit
// tiles), so a wave lets its lanes run dry once per run of one mesh's items
// rather than once per tile; the first item of another mesh or kind ends the
// walk and starts the next one.
struct ItemSlices {
  static constexpr bool kActive = true;
  const BatchArgs& args;
  uint32_t xcd, k;  // segment cursor: segment (xcd + k) % BATCH_XCDS
  uint32_t mesh, kind, nrays;
  BatchWork held;   // the walk's first item, then the item that ended it
  bool first, have_held;
  // The wave's next item, its own XCD's segment first (every lane active).
  __device__ bool take(BatchWork& w) {
    for (; k < (uint32_t)BATCH_XCDS; ++k) {
      const uint32_t q = (xcd + k) % BATCH_XCDS;
      const uint32_t b0 = args.trace_seg[q], b1 = args.trace_seg[q + 1];
      if (b1 == b0) continue;
      uint32_t it = 0;
      if ((threadIdx.x & 63u) == 0) it = b0 + atomicAdd(args.ctl + BATCH_CTL_XTICKET + q, 1u);
      it = __builtin_amdgcn_readfirstlane(it);
      if (it < b1) {
        w = load_uniform(args.items + it);
        return true;
      }
    }
    return false;
  }
  __device__ bool next(uint32_t& b, uint32_t& e) {
    BatchWork w;
    if (first) {
      w = held;
      first = false;
    } else {
      if (!take(w)) return false;
      if (w.mesh != mesh || (w.tile & BATCH_ANY_BIT) != kind) {
        held = w, have_held = true;
        return false;
      }
    }
    b = (w.tile & ~BATCH_ANY_BIT) * args.ray_tile;
    e = min(b + args.ray_tile, nrays);
    return true;
  }
};

// One wave's walk over the rays of one mesh and query kind, its slices from
// src (the whole mesh's arrays: a ray's index is its index in the mesh).
template <int DEPTH, bool TSTACK, bool WIDE>
__device__ inline void trace_mesh(const BatchArgs& a, const BatchMeshDev& m, uint32_t* lds,
                                  ItemSlices* src) {
  TraceArgs t{};
  t.nodes = m.v.B.nodes;
  t.wide = m.v.B.wide;
  t.tris = m.v.B.tris;
  t.nprims = m.v.n;
  t.rays = m.rays;
  t.nrays = m.nrays;
  t.gstack = a.gstack;
  t.gdepth = a.gdepth;
  t.err = a.ctl + BATCH_CTL_TRACE_ERR;
  t.refill_idle = a.refill_idle;
  t.chunk = 64;
  t.sphere = 0;  // batch meshes are triangle scenes
  int32_t* s_node = reinterpret_cast<int32_t*>(lds);
  float* s_tn = reinterpret_cast<float*>(lds + DEPTH * BUILD_BLOCK);
  const uint32_t gtid = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  const uint32_t gstride = gridDim.x * BUILD_BLOCK;
  const RayStream stream;  // unused: the slices come from src
  if (src->kind) {
    t.occ = m.occ;
    trace_loop<true, false, BUILD_BLOCK, DEPTH, false, false, WIDE>(t, stream, s_node, s_tn, gtid,
                                                                    gstride, src);
  } else {
    t.prim = m.prim;
    t.t = m.t;
    t.u = m.u;
    t.v = m.v_;
    trace_loop<false, false, BUILD_BLOCK, DEPTH, TSTACK, false, WIDE>(t, stream, s_node, s_tn,
                                                                      gtid, gstride, src);
  }
}

// The batch build's refit chunks: one leaf per thread of a block. A work
// item of the refit phase covers BATCH_REFIT_CHUNKS chunks, one of the
// spanning-node phases BATCH_SPAN_TILES tiles of BUILD_BLOCK internal nodes:
// every item costs a ticket, an atomic on one word that all blocks share,
// and items of a single chunk or tile made the phases wait on those tickets.
constexpr int BATCH_REFIT_CHUNK = BUILD_BLOCK;

// LDS words of the batch kernel: the build kernel's largest phase (lean sort
// tile, Morton histograms, refit chunk), the trace kernel's stack.
constexpr int cmax(int x, int y) { return x > y ? x : y; }
constexpr int batch_lds_words(bool build, int depth, bool tstack) {
  return build ? cmax(cmax(cmax(sort_lds_words(SORT_ITEMS_LEAN), SORT_PASSES * 256),
                           refit_chunk_lds_words<BATCH_REFIT_CHUNK>()),
                      karras_lds_words<NODE_TILE>())
               : depth * BUILD_BLOCK * (tstack ? 2 : 1);
}

// Two kernels share this body. KM_BUILD runs the build phases of every mesh
// (init ... widen) with grid barriers between them; KM_TRACE runs the trace
// phase alone (ray tiles over built BVHs, no barrier): after KM_BUILD in a
// batch run (WIDE: walking the 4-wide nodes of the widen phase), or as the
// continuous-batching step of the serving loop (binary nodes). WAVES waves
// per SIMD cap the registers: the build phases are written to fit 8 (lean
// 8-key sort tiles, descriptors loaded per phase); the 4-wide walk fits 6
// (79 VGPRs) where its LDS stack allows, the binary walk 7-8.
enum KernelMode { KM_BUILD, KM_TRACE };
template <int DEPTH, bool TSTACK, int WAVES, KernelMode MODE, bool WIDE, bool WAVE_ITEMS = WIDE>
__global__ __launch_bounds__(BUILD_BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_batch(
    BatchArgs a) {
  __shared__ uint32_t lds[batch_lds_words(MODE == KM_BUILD, DEPTH, TSTACK)];
  __shared__ uint32_t s_item, s_wctr[BUILD_BLOCK / 64];
  uint32_t* const ctl = a.ctl;
  const BatchWork* const items = a.items;
  const BatchMeshDev* const meshes = a.meshes;
  if constexpr (MODE == KM_TRACE) {
    // Trace items from the XCD's segment first (the XCD comes from a hardware
    // register; any value is safe: every segment is drained). WAVE_ITEMS
    // (the batch run): waves take items on their own and never wait for the
    // rest of their block; otherwise (the serving step, whose tiles are sized
    // for a block) a block takes an item and its waves share it.
    const uint32_t xcd = __builtin_amdgcn_s_getreg(XCC_ID_HWREG) % BATCH_XCDS;
    if constexpr (WAVE_ITEMS) {
      ItemSlices src{a, xcd, 0u, 0u, 0u, 0u, BatchWork{}, false, false};
      if (!src.take(src.held)) return;
      for (;;) {
        const BatchMeshDev m = load_uniform(meshes + src.held.mesh);
        src.mesh = src.held.mesh, src.kind = src.held.tile & BATCH_ANY_BIT, src.nrays = m.nrays;
        src.first = true, src.have_held = false;
        trace_mesh<DEPTH, TSTACK, WIDE>(a, m, lds, &src);
        if (!src.have_held) return;
      }
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (int k = 0; k < BATCH_XCDS; ++k) {
      const uint32_t q = (xcd + (uint32_t)k) % BATCH_XCDS;
      const uint32_t b0 = a.trace_seg[q], b1 = a.trace_seg[q + 1];
      if (b1 == b0) continue;
      uint32_t* const ticket = ctl + BATCH_CTL_XTICKET + q;
      for (;;) {
        uint32_t it = 0;
        if (WAVE_ITEMS) {
          if (lane == 0) it = b0 + atomicAdd(ticket, 1u);
          it = __builtin_amdgcn_readfirstlane(it);  // every lane is active here
        } else {
          if (threadIdx.x == 0) s_item = b0 + atomicAdd(ticket, 1u);
          __syncthreads();
          it = __builtin_amdgcn_readfirstlane(s_item);
          __syncthreads();
        }
        if (it >= b1) break;
        const BatchWork w = load_uniform(items + it);
        const BatchMeshDev m = load_uniform(meshes + w.mesh);
        trace_item<DEPTH, TSTACK, WIDE, WAVE_ITEMS>(a.gstack, a.gdepth,
                                                    ctl + BATCH_CTL_TRACE_ERR, a.ray_tile,
                                                    a.refill_idle, m, w.tile, lds,
                                                    s_wctr + (WAVE_ITEMS ? wave : 0u));
      }
    }
    return;
  }
  // ---- the build phases, block by block, grid barriers between them
  // blocks the device places after the membership closed take no part
  const uint32_t ticket = join_grid(ctl, &s_item);
  if (ticket == ~0u) return;
  if (threadIdx.x == 0) atomicMin(clock_stamp(ctl, 0), (unsigned long long)wall_clock64());
  constexpr int first = 0, last = (int)PH_TRACE;
  for (int ph = first; ph < last; ++ph) {
   {
    // kernarg words, block-uniform
    const uint32_t b0 = a.phase_begin[ph], b1 = a.phase_begin[ph + 1];
    uint32_t* const ticket = ctl + BATCH_CTL_TICKET + ph;
    for (;;) {
      if (b1 == b0) break;
      if (threadIdx.x == 0) s_item = b0 + atomicAdd(ticket, 1u);
      __syncthreads();
      // block-uniform: keep the item and its mesh descriptor in scalar registers
      const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
      __syncthreads();
      if (it >= b1) break;
      const BatchWork w = load_uniform(items + it);
      const BatchMeshDev* const mp = meshes + w.mesh;
      // each phase loads the descriptor itself: fields it does not read are
      // never loaded and hold no scalar registers
#define VIEW const BuildView v = load_uniform(&mp->v)
      switch (ph) {
        case PH_INIT: {
          VIEW;
          build_init_part(v, threadIdx.x, BUILD_BLOCK);
          break;
        }
        case PH_BOUNDS: {
          VIEW;
          bounds_tile(v, w.tile, lds);
          break;
        }
        case PH_MORTON: {
          VIEW;
          morton_tile(v, w.tile, lds);
          break;
        }
        case PH_SORT0: {
          VIEW;
          onesweep_tile<0, SORT_ITEMS_LEAN>(v, w.tile, lds);
          break;
        }
        case PH_SORT1: {
          VIEW;
          onesweep_tile<1, SORT_ITEMS_LEAN>(v, w.tile, lds);
          break;
        }
        case PH_SORT2: {
          VIEW;
          onesweep_tile<2, SORT_ITEMS_LEAN>(v, w.tile, lds);
          break;
        }
        case PH_SORT3: {
          VIEW;
          onesweep_tile<3, SORT_ITEMS_LEAN>(v, w.tile, lds);
          break;
        }
        case PH_KARRAS: {
          VIEW;
          karras_tile(v, w.tile, lds);
          break;
        }
        case PH_REFIT: {  // chunk-union refit, pass 1 (with the chunks' wide records)
          VIEW;
          for (uint32_t c = 0; c < (uint32_t)BATCH_REFIT_CHUNKS; ++c) {
            const uint32_t chunk = w.tile * BATCH_REFIT_CHUNKS + c;
            if (chunk * BATCH_REFIT_CHUNK >= v.n) break;
            refit_chunk_tile<BATCH_REFIT_CHUNK>(v, chunk, lds);
            __syncthreads();  // the chunk's LDS is read to the end before the next one
          }
          break;
        }
        case PH_TOP: {  // pass 2: the nodes spanning chunks, a wave per 64 nodes
          VIEW;
          for (uint32_t t = 0; t < (uint32_t)BATCH_SPAN_TILES; ++t)
            refit_span_wave<BATCH_REFIT_CHUNK>(
                v.B, v.n, (w.tile * BATCH_SPAN_TILES + t) * BUILD_BLOCK + threadIdx.x);
          break;
        }
        case PH_WIDEN: {  // wide records of the spanning nodes
          VIEW;
          for (uint32_t t = 0; t < (uint32_t)BATCH_SPAN_TILES; ++t) {
            const uint32_t i = (w.tile * BATCH_SPAN_TILES + t) * BUILD_BLOCK + threadIdx.x;
            if (i + 1u < v.n && spans_chunks<BATCH_REFIT_CHUNK>(v.B, i)) widen_global(v.B, i);
          }
          break;
        }
        default:
          break;
      }
#undef VIEW
      __syncthreads();
    }
   }
    if (ph + 1 < last) grid_barrier(ctl, (uint32_t)(ph + 1), ticket);
  }
  if (threadIdx.x == 0) atomicMax(clock_stamp(ctl, last), (unsigned long long)wall_clock64());
}

// Clears the queue tickets, the barrier and the error words before a run (a
// kernel, so that the reset lives in the same captured graph as the run).
__global__ void k_batch_reset(uint32_t* ctl) {
  static_assert(BATCH_CTL_WORDS <= 64 && 2 * BATCH_BAR_GROUPS <= 64,
                "one 64-thread block clears the control and barrier words");
  // the start stamp is a minimum: it starts at the largest value
  const bool start = threadIdx.x == BATCH_CTL_CLOCK || threadIdx.x == BATCH_CTL_CLOCK + 1;
  if (threadIdx.x < BATCH_CTL_WORDS) ctl[threadIdx.x] = start ? 0xFFFFFFFFu : 0u;
  if (threadIdx.x < 2 * BATCH_BAR_GROUPS) *bar_group(ctl, threadIdx.x) = 0u;
}

typedef void (*BatchFn)(BatchArgs);
struct BatchVariant {
  int depth;
  bool tstack;
  BatchFn fn;
};
// The build kernel of a batch run, by waves per SIMD (the register cap trades
// occupancy against spills of the sort phase's look-back window).
const BatchFn kBuild[] = {k_batch<0, false, 8, KM_BUILD, false>, k_batch<0, false, 7, KM_BUILD, false>,
                          k_batch<0, false, 6, KM_BUILD, false>, k_batch<0, false, 5, KM_BUILD, false>,
                          k_batch<0, false, 4, KM_BUILD, false>};
constexpr int kBuildWaves[] = {8, 7, 6, 5, 4};
constexpr int kNumBuildVariants = sizeof(kBuild) / sizeof(kBuild[0]);
// Trace kernels of a batch run (4-wide walk): the batch configurations.
const BatchVariant kBatchVariants[] = {
    {24, false, k_batch<24, false, 6, KM_TRACE, true>},
    {16, false, k_batch<16, false, 6, KM_TRACE, true>},
    {32, false, k_batch<32, false, 4, KM_TRACE, true>},
    {12, true, k_batch<12, true, 6, KM_TRACE, true>},
    {28, false, k_batch<28, false, 5, KM_TRACE, true>},
    {30, false, k_batch<30, false, 5, KM_TRACE, true>}};
// Trace kernels of the serving loop: binary walks (0-5), then 4-wide walks
// with block-shared items (6-7; the slots keep wide nodes).
const BatchVariant kServeVariants[] = {
    {16, false, k_batch<16, false, 7, KM_TRACE, false>}, {20, false, k_batch<20, false, 7, KM_TRACE, false>},
    {8, true, k_batch<8, true, 7, KM_TRACE, false>},     {10, true, k_batch<10, true, 7, KM_TRACE, false>},
    {16, false, k_batch<16, false, 8, KM_TRACE, false>}, {8, true, k_batch<8, true, 8, KM_TRACE, false>},
    {16, false, k_batch<16, false, 5, KM_TRACE, true, false>},
    {16, false, k_batch<16, false, 6, KM_TRACE, true, false>}};
constexpr int kServeWideFirst = 6;
constexpr int kNumBatchVariants = sizeof(kBatchVariants) / sizeof(kBatchVariants[0]);
constexpr int kNumServeVariants = sizeof(kServeVariants) / sizeof(kServeVariants[0]);

uint32_t grid_blocks(int device, BatchFn fn) {
  int cus = 0, per = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, BUILD_BLOCK, 0);
  if (cus <= 0) cus = 256;
  if (per <= 0) per = 1;
  return (uint32_t)(cus * per);
}

}  // namespace

int batch_num_variants() { return kNumBatchVariants; }
void batch_variant_info(int v, int* depth, int* tstack) {
  *depth = kBatchVariants[v].depth;
  *tstack = kBatchVariants[v].tstack ? 1 : 0;
}
int serve_num_variants() { return kNumServeVariants; }
void serve_variant_info(int v, int* depth, int* tstack, int* wide) {
  *depth = kServeVariants[v].depth;
  *tstack = kServeVariants[v].tstack ? 1 : 0;
  *wide = v >= kServeWideFirst ? 1 : 0;
}
int batch_num_build_variants() { return kNumBuildVariants; }
int batch_build_waves(int b) { return kBuildWaves[b]; }
uint32_t batch_build_grid_blocks(int device, int b) { return grid_blocks(device, kBuild[b]); }
uint32_t batch_trace_grid_blocks(int device, int v) {
  return grid_blocks(device, kBatchVariants[v].fn);
}
uint32_t serve_grid_blocks(int device, int v) { return grid_blocks(device, kServeVariants[v].fn); }

hipError_t launch_batch(const LaunchCtx& lc, const BatchArgs& a, int build, uint32_t build_grid,
                        int variant, uint32_t trace_grid) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, lc.stream, a.ctl);
  {
    ProfScope ps(lc, "batch_build");
    hipLaunchKernelGGL(kBuild[build], dim3(build_grid), dim3(BUILD_BLOCK), 0, lc.stream, a);
  }
  {
    ProfScope ps(lc, "batch_trace");
    hipLaunchKernelGGL(kBatchVariants[variant].fn, dim3(trace_grid), dim3(BUILD_BLOCK), 0,
                       lc.stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_serve_trace(const LaunchCtx& lc, const BatchArgs& a, int variant, uint32_t grid) {
  ProfScope ps(lc, "batch_trace");
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, lc.stream, a.ctl);
  hipLaunchKernelGGL(kServeVariants[variant].fn, dim3(grid), dim3(BUILD_BLOCK), 0, lc.stream, a);
  return hipGetLastError();
}

}  // namespace rq
