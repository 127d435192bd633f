// Internal types of the persistent batch kernel (batch_kernels.hip <-> batch.cpp).
#pragma once
#include "build_device.h"

namespace rq {

enum BatchPhase {
  PH_INIT = 0,
  PH_BOUNDS,
  PH_MORTON,
  PH_SORT0,
  PH_SORT1,
  PH_SORT2,
  PH_SORT3,
  PH_KARRAS,
  PH_REFIT,
  PH_TOP,
  PH_WIDEN,  // 4-wide nodes from the refitted binary records
  PH_TRACE,
  BATCH_PHASES
};

// The trace items are split into one queue segment per XCD (each XCD has its
// own L2): a mesh's rays are traced on one XCD so its BVH stays in that L2;
// a block drains its own XCD's segment first, then takes from the others.
constexpr int BATCH_XCDS = 8;

// control words of one batch
constexpr int BATCH_CTL_TICKET = 0;                        // BATCH_PHASES queue tickets
constexpr int BATCH_CTL_BARRIER = BATCH_PHASES;            // grid barrier arrivals
constexpr int BATCH_CTL_BARRIER_ERR = BATCH_PHASES + 1;    // barrier wait gave up (this run)
constexpr int BATCH_CTL_TRACE_ERR = BATCH_PHASES + 2;      // stack deeper than provided (this run)
constexpr int BATCH_CTL_XTICKET = BATCH_PHASES + 3;        // BATCH_XCDS trace segment tickets
constexpr int BATCH_CTL_START = BATCH_CTL_XTICKET + BATCH_XCDS;  // blocks started (+ BATCH_CLOSED)
constexpr int BATCH_CTL_MEMBERS = BATCH_CTL_START + 1;     // blocks taking part (0 until closed)
// BATCH_PHASES 64-bit wall-clock stamps (8-byte aligned) of the build kernel:
// stamp 0 = the first member's start, stamp p = phase p - 1 done (the last
// member through barrier p, or out of the kernel after the last phase)
constexpr int BATCH_CTL_CLOCK = (BATCH_CTL_MEMBERS + 2) & ~1;
constexpr int BATCH_CTL_WORDS = BATCH_CTL_CLOCK + 2 * BATCH_PHASES;
// The grid barrier's group words, after the control words: BATCH_BAR_GROUPS
// arrival counters, then as many release words, each word on a 4 KB page of
// its own, so that the arrivals and polls of different groups are served by
// different memory channels rather than queueing on one word.
constexpr int BATCH_BAR_GROUPS = 16;
constexpr int BATCH_BAR_STRIDE = 1024;  // words
constexpr int BATCH_BAR_BASE = BATCH_BAR_STRIDE;
constexpr int BATCH_CTL_ALLOC_WORDS = BATCH_BAR_BASE + 2 * BATCH_BAR_GROUPS * BATCH_BAR_STRIDE;
static_assert(BATCH_CTL_WORDS <= BATCH_BAR_BASE, "control words before the barrier groups");
constexpr uint32_t BATCH_CLOSED = 0x80000000u;             // start word: membership closed
constexpr uint32_t BATCH_ANY_BIT = 0x80000000u;            // trace item tile: any-hit query
// Work items of the refit phases (see batch_kernels.hip): refit chunks of
// BUILD_BLOCK leaves, BATCH_REFIT_CHUNKS per item; spanning nodes in tiles of
// BUILD_BLOCK, BATCH_SPAN_TILES per item.
constexpr uint32_t BATCH_REFIT_CHUNKS = 4;
constexpr uint32_t BATCH_SPAN_TILES = 8;

struct BatchMeshDev {
  BuildView v;
  rq_rays rays;
  uint32_t nrays;
  int32_t* prim;
  float *t, *u, *v_;
  uint8_t* occ;
};

struct BatchWork {
  uint32_t mesh;
  uint32_t tile;  // unit of the phase; trace: ray tile, BATCH_ANY_BIT for any-hit
};

struct BatchArgs {
  const BatchMeshDev* meshes;
  const BatchWork* items;               // every phase's items, phase after phase
  uint32_t phase_begin[BATCH_PHASES + 1];
  uint32_t trace_seg[BATCH_XCDS + 1];   // trace items of XCD q: [trace_seg[q], trace_seg[q + 1])
  uint32_t* ctl;                        // BATCH_CTL_WORDS
  int32_t* gstack;                      // traversal stack overflow of the persistent grid
  uint32_t gdepth;
  uint32_t ray_tile;
  uint32_t refill_idle;
};

// Batch configurations: the trace kernels of a batch run (4-wide walk).
int batch_num_variants();
void batch_variant_info(int v, int* depth, int* tstack);
// Build kernels by waves per SIMD (8, 7, 6, 5).
int batch_num_build_variants();
int batch_build_waves(int build);
uint32_t batch_build_grid_blocks(int device, int build);
uint32_t batch_trace_grid_blocks(int device, int v);
// The serving loop's trace kernels (binary walk over its slots).
int serve_num_variants();
void serve_variant_info(int v, int* depth, int* tstack, int* wide);
uint32_t serve_grid_blocks(int device, int v);
// A batch run: resets the control words, then the build kernel (every build
// phase, grid barriers between them) and the trace kernel of `variant`.
hipError_t launch_batch(const LaunchCtx& lc, const BatchArgs& a, int build, uint32_t build_grid,
                        int variant, uint32_t trace_grid);
// The serving step: resets the control words, then the trace phase alone.
hipError_t launch_serve_trace(const LaunchCtx& lc, const BatchArgs& a, int variant, uint32_t grid);

}  // namespace rq
