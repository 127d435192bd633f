/*
 * rq — wavefront renderer on top of the ray-query ABI (C ABI).
 *
 * Replaces render() (main.cpp:541-566) + castRay (main.cpp:291-500) for the
 * GPU_LBVH data structure. The reference walks pixels one by one and recurses
 * per ray; here each stage is one batched launch over all samples:
 * raygen -> closest hit -> shade (spawn shadow / bounce rays) -> [closest hit
 * of bounce rays -> shade] -> any-hit of shadow rays -> resolve.
 *
 * Shading is castRay's DIFFUSE_AND_GLOSSY branch (main.cpp:447-492) with a
 * real shadow test in place of the trace_more stub (main.cpp:235-245): the
 * shadow ray starts at the hit point offset by bias along the normal and is
 * occluded if any triangle lies strictly between it and the light. With
 * bounce = 1 a mirror ray reflect(d, N) is traced from every primary hit
 * (origin offset as main.cpp:438-441) and its direct light is added weighted
 * by fresnel(d, N, 2) (main.cpp:85-103, :444); a missed bounce sees the sky.
 */
#ifndef RQ_RENDER_H
#define RQ_RENDER_H
#include <stddef.h>
#include <stdint.h>

#include "rq.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Camera jitter source. The camera is the reference's (main.cpp:544-557):
 * origin 0, fov 30 degrees, direction normalize(xx, yy, -1), a jitter pair
 * (r1, r2) per sample always applied. */
#define RQ_CAMERA_REF_JITTER 0 /* std::mt19937 (default seed) + uniform_real_distribution<double>, main.cpp:503-508 */
#define RQ_CAMERA_PHILOX 1     /* Philox4x32-10 keyed by seed, counter = sample index */

typedef struct {
  uint32_t width, height, spp;
  int32_t camera;
  uint64_t seed;
  float light[3];    /* point light position (main.cpp:775: (0, 3, 30)) */
  float emission[3]; /* light emission (1, 1, 1) */
  float surface[3];  /* surfaceColor added per light (main.cpp:490) */
  float sky[3];      /* miss colour (0.6, 0.8, 1) (main.cpp:312, :341) */
  float bias;        /* 1e-4 (main.cpp:404) */
  int32_t bounce;    /* 0: primary + shadow; 1: + mirror bounce and its shadow ray */
  int32_t shadows;   /* 1: a shadow ray per hit (any hit); 0: the reference's trace_more
                        stub (main.cpp:235-245): lights never occluded, no shadow rays */
  int32_t materials; /* 1: castRay's material switch (main.cpp:417-446) on the scene's
                        materials (rq_scene_set_materials): a REFLECTION_AND_REFRACTION hit
                        traces its refraction ray (refract ior 3, from p -/+ N*bias) and
                        weights its colour by 1 - fresnel(ior 2); REFLECTION gives the grey
                        1 - fresnel; at the second hit a refracting surface shows the sky
                        (depth 3). Not combined with bounce. */
  int32_t mode;      /* RQ_RENDER_SHADE, or RQ_RENDER_OCCLUSION: castRay's KDTREE case
                        (main.cpp:362-370), black where the primary ray hits anything
                        (any hit), else the sky */
} rq_render_params;

#define RQ_RENDER_SHADE 0
#define RQ_RENDER_OCCLUSION 1

typedef struct {
  uint64_t primary, bounce, shadow; /* rays traced in the last frame (active only) */
} rq_frame_counts;

typedef struct rq_renderer rq_renderer;

/* Fills p with the reference defaults (640x480, 1 spp, REF_JITTER, light
 * (0,3,30), emission 1, surface 0, sky (0.6,0.8,1), bias 1e-4, no bounce,
 * shadows on, no materials, RQ_RENDER_SHADE). */
void rq_render_defaults(rq_render_params* p);
int rq_renderer_create(rq_ctx* ctx, const rq_render_params* p, rq_renderer** out);
int rq_renderer_destroy(rq_ctx* ctx, rq_renderer* r);
/* Enqueue one frame on `stream` against bvh (its scene supplies normals). */
int rq_render_frame(rq_ctx* ctx, rq_renderer* r, const rq_bvh* bvh, void* stream);
/* Makes `stream` wait, on the device (no host wait), until the last frame
 * enqueued by rq_render_frame has finished its last closest-hit launch (the
 * any-hit launch of an occlusion frame): the rest of that frame is shading and
 * the second hits' shadow rays. Frames on one stream run in order, so every
 * earlier frame is then complete. A rebuild of the BVH the previous frame read
 * can start there and overlap the frame's tail (rq_stream_order after the
 * whole frame is the other choice). No frame yet: returns at once.
 * No counterpart in the reference (its frames are sequential host loops). */
int rq_render_wait_traced(rq_ctx* ctx, rq_renderer* r, void* stream);
int rq_render_counts(rq_ctx* ctx, rq_renderer* r, rq_frame_counts* out);
/* Image after the last frame: rgb = per-pixel sum over samples (w*h*3
 * floats, may be NULL); ppm = (uint8)(min(1, rgb/spp) * 255) (main.cpp:521-523). */
int rq_render_read_image(rq_ctx* ctx, rq_renderer* r, float* rgb, uint8_t* ppm);
/* Per-sample results of the last frame, any pointer may be NULL:
 * primary/bounce hit ids (-1 = miss or no ray), shadow flags, sample colour (3 floats). */
int rq_render_read_samples(rq_ctx* ctx, rq_renderer* r, int32_t* hit0, int32_t* hit1,
                           uint8_t* shadow0, uint8_t* shadow1, float* sample_rgb);
/* Device SoA of the primary rays of the last frame (valid until destroy).
 * Read-only: the origin, tmin and tmax arrays of the renderer's batches are
 * shared constant arrays (zeros, +inf). */
int rq_render_primary_rays(rq_ctx* ctx, rq_renderer* r, rq_rays* out, uint32_t* n);
/* Device SoA of one ray batch of the last frame: which = 0 primary, 1 bounce,
 * 2 shadow (n or 2n slots; unused slots are inactive rays). Read-only, like
 * the primary rays: every batch's tmin is the renderer's shared zero array. */
int rq_render_rays(rq_ctx* ctx, rq_renderer* r, int which, rq_rays* out, uint32_t* n);

/* Batch ray generators into caller-owned device SoA buffers. */
int rq_raygen_camera(rq_ctx* ctx, const rq_render_params* p, const double* jitter_dev,
                     float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                     float* tmin, float* tmax, void* stream);
/* n rays: origin uniform in [lo, hi], direction uniform on the sphere
 * (rejection from the cube, at most 64 Philox draws, then +z), t in (0, inf). */
int rq_raygen_uniform(rq_ctx* ctx, uint32_t n, uint64_t seed, const float lo[3], const float hi[3],
                      float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                      float* tmin, float* tmax, void* stream);

/* P6 PPM writer (main.cpp:518-524 layout). */
int rq_write_ppm(const char* path, uint32_t w, uint32_t h, const uint8_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
