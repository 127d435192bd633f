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
constexpr int BATCH_CTL_WORDS = BATCH_CTL_MEMBERS + 1;
constexpr uint32_t BATCH_CLOSED = 0x80000000u;             // start word: membership closed
constexpr uint32_t BATCH_ANY_BIT = 0x80000000u;            // trace item tile: any-hit query

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
void serve_variant_info(int v, int* depth, int* tstack);
uint32_t serve_grid_blocks(int device, int v);
// A batch run: resets the control words, then the build kernel (every build
// phase, grid barriers between them) and the trace kernel of `variant`.
hipError_t launch_batch(const LaunchCtx& lc, const BatchArgs& a, int build, uint32_t build_grid,
                        int variant, uint32_t trace_grid);
// The serving step: resets the control words, then the trace phase alone.
hipError_t launch_serve_trace(const LaunchCtx& lc, const BatchArgs& a, int variant, uint32_t grid);

}  // namespace rq
