/*
 * rq — serving: a request loop over resident BVHs with continuous batching.
 *
 * The reference renders one scene per process run (main.cpp:723-853) and has
 * no request path. Here a server owns one device: a pool of resident mesh
 * slots (vertex buffers + LBVH, built on the device), an arena for request
 * buffers, and a host thread that runs the loop:
 *
 *   submit (any thread)  -> queued
 *   admission (FIFO)     -> a request is admitted when an in-flight slot is
 *                           free and its ray + result buffers fit the arena
 *                           (and, when it uploads a new mesh, a mesh slot can
 *                           be recycled); its rays are generated or copied in
 *                           and a new mesh is built, on a preparation stream
 *   continuous batching  -> every step is ONE trace launch over a work queue
 *                           that holds ray tiles of every ready request (up
 *                           to a per-request chunk and a per-step budget);
 *                           requests that became ready while a step ran join
 *                           the next step, and a request larger than its
 *                           chunk continues over several steps
 *   completion           -> results are copied to the caller's host arrays on
 *                           an output stream; the request's buffers and slot
 *                           are released and the next queued request admitted
 *
 * Preparation, traversal and result copies run on three streams, so the
 * uploads and builds of new requests overlap the traversal of admitted ones.
 * The results of a request are identical whether it runs alone or in a mix
 * (traversal is deterministic per ray), and identical to rq_closest_hit /
 * rq_any_hit on the same BVH and rays.
 */
#ifndef RQ_SERVE_H
#define RQ_SERVE_H
#include <stddef.h>
#include <stdint.h>

#include "rq.h"
#include "rq_scene.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct rq_server rq_server;

enum { RQ_RAYS_HOST = 0, RQ_RAYS_PINHOLE = 1, RQ_RAYS_UNIFORM = 2 };
enum { RQ_Q_CLOSEST = 1, RQ_Q_ANY = 2 };
enum { RQ_REQ_QUEUED = 0, RQ_REQ_ADMITTED = 1, RQ_REQ_DONE = 2 };

/* Pinhole camera: ray i goes through pixel (i % width, i / width) of a
 * width x ceil(n / width) image, jittered inside the pixel by Philox(seed, i). */
typedef struct {
  float eye[3], look_at[3], up[3];
  float fov_deg; /* vertical field of view */
  uint32_t width;
} rq_pinhole;

typedef struct {
  uint32_t mesh;           /* resident mesh slot (rq_server_load_mesh); ignored when
                              new_mesh is set                                        */
  const rq_mesh* new_mesh; /* non-NULL: upload this host mesh into a slot that holds
                              no resident mesh, build its LBVH, then trace (host
                              arrays valid until done); the mesh serves this request
                              only, and its slot is recycled by later uploads        */
  uint32_t nrays;
  int32_t source;          /* RQ_RAYS_*                                              */
  rq_rays host_rays;       /* RQ_RAYS_HOST: host SoA arrays, valid until done        */
  rq_pinhole pinhole;      /* RQ_RAYS_PINHOLE                                        */
  float lo[3], hi[3];      /* RQ_RAYS_UNIFORM: origins uniform in [lo, hi]           */
  uint64_t seed;           /* generated rays                                         */
  int32_t query;           /* RQ_Q_CLOSEST | RQ_Q_ANY                                */
  int32_t* prim;           /* host outputs of the closest-hit query (prim required)  */
  float* t;                /*   (t may be NULL)                                      */
  uint8_t* occluded;       /* host output of the any-hit query                       */
} rq_request;

typedef struct {
  int32_t state;       /* RQ_REQ_*, or a negative RQ_E_* code if the request failed */
  uint32_t mesh;       /* slot the request traced (the new mesh's slot for uploads) */
  uint32_t steps;      /* batched trace launches the request took part in           */
  double t_submit;     /* rq_server_clock() times                                   */
  double t_admit;
  double t_done;       /* results are in the host arrays                            */
} rq_request_status;

typedef struct {
  size_t hbm_budget;       /* bytes: mesh slots + arena + traversal scratch          */
  uint32_t mesh_slots;     /* resident BVH slots                                     */
  uint32_t slot_tris;      /* triangle capacity of each slot                         */
  size_t arena_bytes;      /* request buffers (41 B per ray at most) for admission   */
  uint32_t max_inflight;   /* requests admitted at once                              */
  uint32_t step_rays;      /* rays per batched trace launch                          */
  uint32_t chunk_rays;     /* rays of one request per launch                         */
  uint32_t ray_tile;       /* rays per work item of the trace launch                 */
} rq_server_config;

typedef struct {
  uint64_t submitted, completed, failed;
  uint64_t steps;          /* batched trace launches                                 */
  uint64_t rays;           /* rays traced (closest and any counted separately)       */
  uint64_t builds;         /* meshes built                                           */
  uint32_t max_step_requests; /* most requests in one launch                         */
  uint32_t resident;       /* mesh slots holding a mesh                              */
  size_t arena_peak;       /* most arena bytes in use at once                        */
  double busy_s;           /* summed duration of the trace launches (events)         */
} rq_server_stats;

/* Fills the defaults: 64 slots of 2^21 triangles, 8 GiB arena, 64 in flight,
 * 2^22 rays per step, 2^20 per request per step, 4096-ray tiles, 200 GB. */
void rq_server_defaults(rq_server_config* cfg);
/* Allocates everything up front (slots, arena, staging) against the budget
 * and starts the loop thread. */
int rq_server_create(int device, const rq_server_config* cfg, rq_server** out);
/* Stops the loop after the requests in flight and frees everything. */
int rq_server_destroy(rq_server* s);
/* Blocking: uploads and builds a resident mesh into a slot that holds none
 * (a free one, else the least recently used idle upload slot); returns the
 * slot. A resident mesh keeps its slot until the server is destroyed, so the
 * slot id always names it; RQ_E_NOMEM when every slot holds a resident mesh.
 * Uploads (requests with new_mesh) need a slot beyond the resident ones. */
int rq_server_load_mesh(rq_server* s, const rq_mesh* m, uint32_t* slot);
/* Non-blocking, any thread: queues a request. RQ_E_NOMEM if it can never fit
 * (buffers beyond the arena, mesh beyond the slot capacity, an upload when
 * every slot holds a resident mesh); RQ_E_INVALID if `mesh` is not a
 * resident slot. */
int rq_server_submit(rq_server* s, const rq_request* r, uint64_t* ticket);
/* Any thread: asks a queued or running request to stop. It finishes with
 * state RQ_E_CANCELED (its host outputs are left unspecified): at once if it
 * was not admitted yet, else once the trace launches that hold its rays have
 * retired and the result copies of the chunks they traced are over (no
 * further chunk is traced); the launches of other requests are not
 * affected. RQ_E_INVALID for an unknown ticket or a finished request. */
int rq_server_cancel(rq_server* s, uint64_t ticket);
int rq_server_poll(rq_server* s, uint64_t ticket, rq_request_status* st);
/* Blocks until the request is done or failed. */
int rq_server_wait(rq_server* s, uint64_t ticket, rq_request_status* st);
/* Up to `cap` requests finished since the last call (done or failed), in
 * completion order: their tickets and statuses; *count is the number taken. */
int rq_server_take_done(rq_server* s, uint64_t* tickets, rq_request_status* st, uint32_t cap,
                        uint32_t* count);
/* Forgets a finished request's record. */
int rq_server_forget(rq_server* s, uint64_t ticket);
int rq_server_stats_get(rq_server* s, rq_server_stats* st);
const char* rq_server_last_error(rq_server* s);
/* Seconds on the server's monotonic clock (the clock of rq_request_status). */
double rq_server_clock(void);
/* The rays of an RQ_RAYS_PINHOLE request, into caller device arrays (for
 * checking a request against rq_closest_hit / rq_any_hit on the same rays). */
int rq_raygen_pinhole(rq_ctx* ctx, const rq_pinhole* cam, uint32_t n, uint64_t seed, float* ox,
                      float* oy, float* oz, float* dx, float* dy, float* dz, float* tmin,
                      float* tmax, void* stream);
/* Page-locked host memory for request arrays (fast asynchronous copies). */
void* rq_host_alloc(size_t bytes);
void rq_host_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
