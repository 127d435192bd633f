/*
 * rq — MI355X-native ray-query engine: C ABI.
 *
 * Drop-in GPU_LBVH backend for the acceleration-structure path of
 * Alhajras/Raytracer-Data-structures. The reference has no plugin interface;
 * its boundary is two switch statements over free functions (build side
 * main.cpp:791-843, query side main.cpp:330-388). Each entry point below names
 * the reference function it replaces.
 *
 * Conventions
 *  - Every function returns RQ_OK (0) or a negative RQ_E_* code; nothing
 *    throws across the ABI. rq_last_error() describes the last failure.
 *  - Host arrays belong to the caller. Device memory behind opaque handles
 *    belongs to the library and is charged against the context's HBM budget;
 *    exceeding it returns RQ_E_NOMEM (admission failure) without allocating.
 *  - Ray and hit buffers passed to queries are DEVICE pointers (SoA).
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream); calls
 *    are asynchronous on it and rq_sync() waits. Query launches take one of
 *    the context's 64 ray-stream counters in turn: at most 64 query launches
 *    of one context may be in flight at once, whatever their streams.
 *  - One rq_ctx is used by one host thread at a time.
 */
#ifndef RQ_RQ_H
#define RQ_RQ_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define RQ_OK 0
#define RQ_E_INVALID (-1)     /* bad argument or size                        */
#define RQ_E_NOMEM (-2)       /* HBM budget exceeded or allocation failed    */
#define RQ_E_HIP (-3)         /* HIP runtime error                           */
#define RQ_E_IO (-4)          /* file could not be read or parsed            */
#define RQ_E_NODEVICE (-5)    /* no usable gfx950 device                     */
#define RQ_E_UNSUPPORTED (-6) /* feature not available for this input       */
#define RQ_E_DEVICE (-7)      /* a kernel reported an error (stack, barrier)  */
#define RQ_E_CANCELED (-8)    /* a served request was cancelled              */

/* Morton normalisation (accelerators.h:577 vs Karras 2012). */
#define RQ_MORTON_REF_AFFINE 0 /* (centroid + 30) / 1000, bit-exact with the reference */
#define RQ_MORTON_SCENE_AABB 1 /* (centroid - lo) / (hi - lo) over centroid bounds     */

#define RQ_PRIM_TRIANGLE 0
#define RQ_PRIM_SPHERE 1 /* the reference's primitive (accelerators.h:54-92) */

typedef struct rq_ctx rq_ctx;
typedef struct rq_scene rq_scene;
typedef struct rq_bvh rq_bvh;

typedef struct {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} rq_rays; /* device SoA; a hit needs t in (tmin, tmax); tmin >= tmax = inactive */

typedef struct {
  int32_t* prim; /* primitive index, -1 on miss          */
  float* t;      /* hit distance (+inf on miss), may be NULL */
  float* u;      /* barycentrics, may be NULL            */
  float* v;
} rq_hits; /* device SoA */

typedef struct {
  uint32_t n_prims;    /* leaves                               */
  uint32_t n_internal; /* n_prims - 1 (0 when n_prims <= 1)   */
  uint32_t height;     /* longest root-to-leaf edge count      */
  int32_t morton_mode;
  float root_lo[3], root_hi[3];
  size_t device_bytes; /* HBM held by the BVH                  */
} rq_bvh_info;

/* ---- context ---------------------------------------------------------- */
int rq_create(rq_ctx** ctx, int device, size_t hbm_budget_bytes);
int rq_destroy(rq_ctx* ctx);
const char* rq_last_error(const rq_ctx* ctx);
const char* rq_version(void);
size_t rq_bytes_in_use(const rq_ctx* ctx);

/* ---- scene -------------------------------------------------------------
 * Replaces createScene_new (main.cpp:599-721) + the per-object SceneObject
 * vector: host SoA vertices and a triangle index list (3 per triangle, 0-based)
 * are copied into device-resident SoA buffers. */
int rq_scene_upload(rq_ctx* ctx, const float* vx, const float* vy, const float* vz, uint32_t nv,
                    const uint32_t* idx, uint32_t ntri, int prim_kind, rq_scene** out,
                    void* stream);
/* Spheres (the reference's scene objects, main.cpp:677-689): centres, radii
 * and an optional surface colour per sphere (3 floats each, added by the
 * renderer's shading as sphere->surfaceColor at main.cpp:490; NULL uses the
 * render parameters' colour). Built and queried like triangle scenes; a
 * sphere hit is the reference's geometric test (accelerators.h:79-92) with
 * castRay's far-root rule (main.cpp:349-353), t in [tmin, tmax). */
int rq_scene_upload_spheres(rq_ctx* ctx, const float* cx, const float* cy, const float* cz,
                            const float* radius, const float* rgb, uint32_t n, rq_scene** out,
                            void* stream);
/* Material per primitive for the renderer (`mat` host array of the scene's
 * primitive count), values of the reference's MaterialType (accelerators.h:24):
 * RQ_MAT_DIFFUSE_AND_GLOSSY, RQ_MAT_REFLECTION_AND_REFRACTION, RQ_MAT_REFLECTION.
 * Read by a renderer created with `materials = 1`; a scene without
 * materials is all DIFFUSE_AND_GLOSSY. Blocking; may be called again. */
#define RQ_MAT_DIFFUSE_AND_GLOSSY 0
#define RQ_MAT_REFLECTION_AND_REFRACTION 1
#define RQ_MAT_REFLECTION 2
int rq_scene_set_materials(rq_ctx* ctx, rq_scene* s, const uint8_t* mat, void* stream);
/* Replace the vertex positions of an uploaded scene (same counts); the next
 * rq_rebuild_lbvh sees them. */
int rq_scene_update_vertices(rq_ctx* ctx, rq_scene* s, const float* vx, const float* vy,
                             const float* vz, void* stream);

/* ---- build -------------------------------------------------------------
 * Replaces constructLBVHTree (accelerators.h:570-586) and constructBVHNew
 * (accelerators.h:246-337): Morton codes -> 8-bit-digit radix sort ->
 * Karras internal-node emission -> bottom-up atomic AABB refit, all on the GPU. */
int rq_build_lbvh(rq_ctx* ctx, const rq_scene* s, int morton_mode, rq_bvh** out, void* stream);
/* Rebuild in place from the scene's current vertices (no allocation). */
int rq_rebuild_lbvh(rq_ctx* ctx, rq_bvh* b, void* stream);
/* Optional tree optimisation, no counterpart in the reference (it never
 * optimises a tree; SAH-quality trees are what the LBVH trades for build
 * speed): `rounds` (1..8) rounds of treelet restructuring (Karras & Aila,
 * HPG 2013) of a built BVH in place. Each round reshapes every 7-leaf treelet
 * whose optimal arrangement has a strictly lower sum of node surface areas
 * and no greater height, so queries return the same answers with fewer node
 * visits; the tree height never grows. Waits for the device (it reads the
 * tree height), then runs on `stream`; rq_rebuild_lbvh returns the BVH to
 * the plain LBVH. */
int rq_bvh_treelets(rq_ctx* ctx, rq_bvh* b, int rounds, void* stream);
int rq_bvh_info_get(rq_ctx* ctx, const rq_bvh* b, rq_bvh_info* info);
/* Download the build for parity checks (host arrays, any may be NULL):
 * codes[n] in primitive order, sorted_idx[n], left/right[n-1] (child >= 0 is an
 * internal node, child < 0 is leaf ~child), boxes[(2n-1)*6] internal nodes then
 * leaves, (lo xyz, hi xyz). */
int rq_bvh_export(rq_ctx* ctx, const rq_bvh* b, uint32_t* codes, uint32_t* sorted_idx,
                  int32_t* left, int32_t* right, float* boxes);
/* Download the 4-wide nodes (see rq_trace_variant): for binary internal node
 * i, boxes[i*24 + slot*6 + (lo xyz, hi xyz)] and links[i*4 + slot] (>= 0 wide
 * node, < 0 leaf ~k, 0 unused slot). Every internal node has one; the wide
 * walk reaches the root and the nodes at even depth.
 * RQ_E_UNSUPPORTED for a BVH without wide nodes. */
int rq_bvh_export_wide(rq_ctx* ctx, const rq_bvh* b, float* boxes, int32_t* links);

/* ---- cached BVH blob ----------------------------------------------------
 * The reference rebuilds its tree on every run (main.cpp:791-843). A built
 * LBVH can instead be saved as bytes (a 64-byte header with the primitive
 * kind, Morton mode, counts, a fingerprint of the scene's arrays and a
 * checksum, then codes, sorted indices, parent links, nodes, leaf records and
 * the root box) and loaded into a new BVH of the same scene without a build.
 * Load errors: RQ_E_IO for a damaged or foreign blob, RQ_E_INVALID when the
 * scene does not match the fingerprint. A checksum detects accidents only,
 * so before anything reaches the device the payload is checked against the
 * scene's primitives, in O(n) on the host: every node and leaf reached once
 * from the root through in-range links, parent links and height consistent
 * (the height within the traversal stack), the sorted order a permutation,
 * every leaf record the build's record of its primitive, every child box the
 * exact union below it; any failure is RQ_E_IO. */
int rq_bvh_blob_size(rq_ctx* ctx, const rq_bvh* b, size_t* bytes);
int rq_bvh_save(rq_ctx* ctx, const rq_bvh* b, void* blob, size_t capacity, size_t* written);
int rq_bvh_load(rq_ctx* ctx, const rq_scene* s, const void* blob, size_t bytes, rq_bvh** out,
                void* stream);
int rq_bvh_save_file(rq_ctx* ctx, const rq_bvh* b, const char* path);
int rq_bvh_load_file(rq_ctx* ctx, const rq_scene* s, const char* path, rq_bvh** out);
/* The check rq_bvh_load makes, on host arrays and without a device: a blob
 * against triangles (vx/vy/vz with nv vertices, idx 3 per primitive) or
 * spheres (centres vx/vy/vz, radius; nv == nprim). RQ_OK, RQ_E_IO (damaged,
 * or not a well-formed LBVH of these primitives) or RQ_E_INVALID (a blob of
 * other primitives, or bad arguments); the reason goes to msg (may be NULL). */
int rq_bvh_blob_validate(const void* blob, size_t bytes, int prim_kind, const float* vx,
                         const float* vy, const float* vz, uint32_t nv, const uint32_t* idx,
                         const float* radius, uint32_t nprim, char* msg, size_t msg_cap);

/* ---- queries -----------------------------------------------------------
 * Closest hit replaces boxIntersect (accelerators.h:668-690) plus the
 * candidate loop of castRay (main.cpp:343-358); the result equals the NONE
 * brute-force loop (main.cpp:373-387): smallest t, lowest index on ties.
 * Any-hit replaces kdtreeIntersect (accelerators.h:997-1086): occluded[i] = 1
 * iff some primitive has t in (tmin, tmax). */
int rq_closest_hit(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n, rq_hits* hits,
                   void* stream);
int rq_any_hit(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n, uint8_t* occluded,
               void* stream);
/* rq_closest_hit that allocates nothing. It has no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture on a stream that never queried, tight serving loops, harnesses that
 * forbid hipMalloc). Its scratch (the ray-stream counter and the traversal
 * stack overflow area) is a static device array of 64 MiB per device, held
 * from library load until the process ends, instead of the per-stream area a
 * first query on a stream allocates; the launch grid is capped so that every
 * thread's overflow entries fit (any ray count is accepted; large launches
 * run on fewer waves than rq_closest_hit). The counter is zeroed by a kernel
 * at the start of each call. Every call on one device shares the array:
 * calls that overlap on a device (other streams, other contexts) race and
 * must be ordered by the caller. Results equal rq_closest_hit's. */
int rq_closest_hit_noalloc(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                           rq_hits* hits, void* stream);
/* Host-pointer forms, for callers whose rays and results live in host memory
 * (no counterpart of their own in the reference, whose queries take host
 * values: they let its per-pixel loop call GPU_LBVH unchanged). Synchronous:
 * the rays are staged through a pinned buffer and a device buffer the context
 * keeps (grown on demand, at most 48 MiB each, the device one charged to the
 * budget), in launches of at most 2^20 rays, and the call returns once the
 * results are in the caller's arrays. `rays`/`hits`/`occluded` point to HOST
 * arrays; hits->t/u/v may be NULL. Results equal rq_closest_hit's and
 * rq_any_hit's. Stack-depth errors are reported by rq_device_errors as for
 * the device queries. */
int rq_closest_hit_host(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                        rq_hits* hits, void* stream);
int rq_any_hit_host(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                    uint8_t* occluded, void* stream);
/* One ray from host values, the form of the reference's per-ray calls:
 * rq_cast_ray replaces boxIntersect(rayorig, raydir, tree, boxes) plus the
 * candidate loop of castRay (main.cpp:333-358): the closest primitive with t
 * in (tmin, tmax) (-1 on a miss; t = +inf then), lowest index on ties; uv
 * (2 floats, may be NULL) receives the barycentrics. rq_occluded_ray replaces
 * kdtreeIntersect(rayorig, raydir) (main.cpp:362-366, accelerators.h:997):
 * 1 iff some primitive has t in (tmin, tmax). Each call is a launch and a
 * host wait (tens of microseconds): a loop over many rays is faster through
 * rq_closest_hit_host on arrays. */
int rq_cast_ray(rq_ctx* ctx, const rq_bvh* b, const float orig[3], const float dir[3], float tmin,
                float tmax, int32_t* prim, float* t, float* uv);
int rq_occluded_ray(rq_ctx* ctx, const rq_bvh* b, const float orig[3], const float dir[3],
                    float tmin, float tmax, uint8_t* occluded);
/* Instrumented variants: per-ray node-record visits and triangle tests of the
 * canonical traversal (device uint32 arrays). Slower; for measurement. */
int rq_closest_hit_counted(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                           rq_hits* hits, uint32_t* visits, uint32_t* tests, void* stream);
int rq_any_hit_counted(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                       uint8_t* occluded, uint32_t* visits, uint32_t* tests, void* stream);

/* ---- traversal kernel variants and the autotuner ----------------------
 * Variants differ in threads per block, LDS stack entries per lane, whether
 * stack entries carry their entry distance (closest hit skips entries
 * beyond the best hit when popped), and the node format walked: binary
 * (64-byte nodes, two children) or wide (the binary tree collapsed into
 * 128-byte 4-child nodes, written by every rq_build_lbvh / rq_rebuild_lbvh /
 * rq_bvh_load), and the waves per SIMD the closest-hit and any-hit kernels
 * are compiled for (0: as many as their registers allow; 6: registers
 * capped for six, which hides more of the dependent node fetches). Results
 * are identical for all. A new context uses wide variants (9 for closest
 * hit, 14 for any hit, refill 64, slices of 1024). */
typedef struct {
  int32_t block, lds_depth, tstack, wide;
  int32_t closest_waves, any_waves;
} rq_trace_variant;
int rq_trace_variant_count(void);
int rq_trace_variant_get(int index, rq_trace_variant* out);
int rq_trace_config_set(rq_ctx* ctx, int closest_variant, int any_variant, uint32_t refill_idle);
int rq_trace_config_get(rq_ctx* ctx, int* closest_variant, int* any_variant,
                        uint32_t* refill_idle);
/* Refill thresholds per query kind: a wave takes new rays once this many of
 * its 64 lanes are idle (64: only when all are; the default). config_set sets
 * both to its refill_idle, config_get reports the closest-hit one. */
int rq_trace_refill_set(rq_ctx* ctx, uint32_t closest_refill, uint32_t any_refill);
int rq_trace_refill_get(rq_ctx* ctx, uint32_t* closest_refill, uint32_t* any_refill);
/* Largest wave slice per query kind: the persistent waves take rays from the
 * launch's ray stream in slices of this many (a power of two, 64..4096;
 * default 1024), halved for launches too small to give every wave two, and a
 * quarter as many over the stream's last quarter. Larger slices cost fewer
 * atomics, smaller ones keep the rays in flight closer together. */
int rq_trace_slice_set(rq_ctx* ctx, uint32_t closest_slice, uint32_t any_slice);
int rq_trace_slice_get(rq_ctx* ctx, uint32_t* closest_slice, uint32_t* any_slice);
/* Times every variant (and refill threshold 16/48/56/64 idle lanes) on the given
 * device rays (default stream, `repeats` launches each after one warm-up), then
 * slice sizes 256/512/2048 against the default 1024 for the best of them, keeps
 * the fastest for that query kind (variant, refill, slice) in the context and
 * reports it (the slice through rq_trace_slice_get). */
int rq_autotune_trace(rq_ctx* ctx, const rq_bvh* b, const rq_rays* rays, uint32_t n, int any,
                      int repeats, int* best_variant, uint32_t* best_refill, double* best_ms);

/* ---- device buffers (for callers without their own allocator) ---------- */
int rq_device_alloc(rq_ctx* ctx, size_t bytes, void** dptr);
int rq_device_free(rq_ctx* ctx, void* dptr);
int rq_memcpy_h2d(rq_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
int rq_memcpy_d2h(rq_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
/* Device-to-device copy by a library kernel (16-byte aligned pointers/size). */
int rq_memcpy_d2d(rq_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
int rq_memset_d(rq_ctx* ctx, void* dst, int value, size_t bytes, void* stream);
/* Waits for the work enqueued on `stream`; NULL: for all work on the device. */
int rq_sync(rq_ctx* ctx, void* stream);
/* Reads and clears the device error bits of past queries (synchronises the
 * device): bit 0 = a ray needed a deeper traversal stack than provided. */
int rq_device_errors(rq_ctx* ctx, uint32_t* bits);
/* Streams for concurrent queries. Queries on different streams may run at
 * once: each non-null stream gets its own traversal-stack overflow area on
 * its first query. Areas are kept for at most 16 streams per context; the
 * 17th stream takes over the area of a stream whose work has all completed
 * (RQ_E_NOMEM only when 16 streams are busy at once). rq_stream_destroy
 * waits for the stream, frees its area and destroys it. A caller's own stream
 * keeps its area until rq_stream_forget or rq_destroy. Queries on one stream
 * run in order and share its area.
 * Graph capture: a query enqueued on a stream under capture zeroes a counter
 * of its own inside the graph, so the graph can be replayed any number of
 * times (at most 64 captured queries of one context may run at once). The
 * stream must have queried once before the capture (its overflow area is
 * allocated then, and allocation is not allowed under capture:
 * RQ_E_UNSUPPORTED otherwise); a graph keeps that area, so it must not run
 * at the same time as other queries on the stream it was captured on. */
int rq_stream_create(rq_ctx* ctx, void** stream);
int rq_stream_destroy(rq_ctx* ctx, void* stream);
/* The context forgets a caller-owned stream: waits for it and frees its
 * overflow area (call before destroying a stream that has queried). */
int rq_stream_forget(rq_ctx* ctx, void* stream);
/* Orders `waiter` after the work enqueued so far on `signaler` (an event
 * recorded on signaler that waiter waits for, no host wait): e.g. the render
 * of a frame after its BVH's rebuild on another stream, and the next rebuild
 * of that BVH after the render. */
int rq_stream_order(rq_ctx* ctx, void* waiter, void* signaler);

/* ---- per-kernel timing (HIP events on the launch stream) ---------------
 * When enabled, every kernel the library launches is bracketed by events on
 * its own stream; rq_profile_read sums the durations per kernel name. */
int rq_profile_enable(rq_ctx* ctx, int on);
/* Times only the kernels named in a comma-separated list (NULL or "": all);
 * the others run without events around them. */
int rq_profile_only(rq_ctx* ctx, const char* names);
int rq_profile_reset(rq_ctx* ctx);
/* Returns the number of kernel names; fills up to cap entries. */
int rq_profile_read(rq_ctx* ctx, char names[][48], double* total_ms, uint64_t* launches, int cap);

/* ---- handles ---------------------------------------------------------- */
int rq_release(rq_ctx* ctx, void* handle); /* rq_scene* or rq_bvh* */

#ifdef __cplusplus
}
#endif
#endif
