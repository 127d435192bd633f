/*
 * rq — batched build and traversal of many meshes in one persistent launch.
 *
 * The reference builds and renders one scene per process run (main.cpp:723-853).
 * Here a batch holds many uploaded meshes of uneven sizes together with one
 * device ray batch per mesh. rq_batch_run builds every mesh's LBVH (Morton,
 * radix sort, Karras, refit) and traces every ray batch in ONE persistent
 * kernel: its blocks pull work items (mesh, phase, tile) from a device work
 * queue, phase by phase across all meshes, with grid-wide barriers between
 * phases. rq_batch_run_graph captures the launch (and its counter reset) in a
 * hipGraph once and replays it. The results of every mesh are identical to
 * rq_build_lbvh + rq_closest_hit / rq_any_hit on that mesh alone.
 *
 * Co-residency: the grid is sized from the kernel's occupancy (blocks per CU
 * x CUs), but the phase barriers never assume that every block is placed:
 * they count only the blocks that have started. The first barrier waits
 * briefly for the rest of the grid, then closes the membership; blocks the
 * device places later (behind kernels of other streams that held CUs) leave
 * without work. Work items come from queues, so the members complete every
 * phase: a contended run is slower, never wrong. rq_batch_occupancy reports
 * how many blocks took part in the last run.
 */
#ifndef RQ_BATCH_H
#define RQ_BATCH_H
#include <stddef.h>
#include <stdint.h>

#include "rq.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const rq_scene* scene; /* uploaded mesh (rq_scene_upload)                       */
  uint32_t nrays;        /* rays of this mesh                                      */
  rq_rays rays;          /* device SoA rays                                        */
  rq_hits hits;          /* device closest-hit outputs (prim required, t/u/v may be NULL);
                            hits.prim == NULL skips the closest-hit query           */
  uint8_t* occluded;     /* device any-hit outputs, NULL skips the any-hit query   */
} rq_batch_item;

typedef struct {
  int32_t lds_depth; /* stack entries per lane in LDS (rest in global memory) */
  int32_t tstack;    /* stack entries carry their entry distance             */
  uint32_t ray_tile; /* rays per work item (at least 256)                    */
  uint32_t refill_idle;
} rq_batch_config;

typedef struct rq_batch rq_batch;

/* Allocates every mesh's BVH and the work queue (charged to the HBM budget).
 * The items' scenes and ray / hit buffers must stay valid until destroy. */
int rq_batch_create(rq_ctx* ctx, const rq_batch_item* items, uint32_t count, int morton_mode,
                    rq_batch** out);
int rq_batch_destroy(rq_ctx* ctx, rq_batch* b);
/* One run: the build kernel (every mesh's build phases and its 4-wide nodes),
 * then the trace kernel (every ray batch). Asynchronous on `stream`. Runs of
 * one batch share its control words: they must not overlap (enqueue them on
 * one stream, or order them). */
int rq_batch_run(rq_ctx* ctx, rq_batch* b, void* stream);
/* Same work replayed from a hipGraph captured on the first call, launched on
 * `stream` (NULL: the null stream, and the call waits for the run). */
int rq_batch_run_graph(rq_ctx* ctx, rq_batch* b, void* stream);
/* Non-owning handle on mesh i's BVH (rq_bvh_export / rq_bvh_info_get / queries). */
int rq_batch_bvh(rq_ctx* ctx, rq_batch* b, uint32_t i, rq_bvh** out);
int rq_batch_config_count(void);
int rq_batch_config_get(int index, rq_batch_config* out);
int rq_batch_set_config(rq_ctx* ctx, rq_batch* b, int index, uint32_t ray_tile,
                        uint32_t refill_idle);
int rq_batch_get_config(rq_ctx* ctx, const rq_batch* b, int* index, uint32_t* ray_tile,
                        uint32_t* refill_idle);
/* The build kernel's waves per SIMD (8, 7, 6, 5 or 4: its register cap, traded
 * against occupancy); the autotuner chooses it after the trace configuration. */
int rq_batch_set_build_waves(rq_ctx* ctx, rq_batch* b, int waves);
int rq_batch_get_build_waves(rq_ctx* ctx, const rq_batch* b, int* waves);
/* Times the trace configurations, ray-tile sizes and refill thresholds on the
 * batch itself (`repeats` runs each), then the build kernels; keeps and
 * reports the fastest. */
int rq_batch_autotune(rq_ctx* ctx, rq_batch* b, int repeats, int* best_index,
                      uint32_t* best_ray_tile, uint32_t* best_refill, double* best_ms);
/* Items of the work queue per phase (11 phases), for inspection. */
int rq_batch_phase_items(rq_ctx* ctx, rq_batch* b, uint32_t* counts, int cap);
/* Synchronises and reports the last run's build phases in ms (the first
 * `cap` of the 10 build phases, init ... widen), from wall-clock stamps the
 * build kernel takes at its grid barriers; returns the number of build
 * phases. A phase the run did not stamp reads -1. */
int rq_batch_phase_ms(rq_ctx* ctx, rq_batch* b, double* ms, int cap);
/* Synchronises and reports the error bits of the last run: 1 = a grid
 * barrier gave up waiting (a member block did not arrive within seconds; the
 * results of that run are invalid), 2 = a ray needed a deeper stack than
 * provided, 4 = a mesh build reported an error. */
int rq_batch_errors(rq_ctx* ctx, rq_batch* b, uint32_t* bits);
/* Synchronises and reports the blocks that took part in the last run
 * (members) and the launch grid: members < grid when part of the grid could
 * not be placed before the first barrier (the device was busy). */
int rq_batch_occupancy(rq_ctx* ctx, rq_batch* b, uint32_t* members, uint32_t* grid);

#ifdef __cplusplus
}
#endif
#endif
