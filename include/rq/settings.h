/*
 * rq — the reference's compile-time configuration with the GPU_LBVH option.
 *
 * Mirrors settings.h:7-17 and the AccType enum of accelerators.h:21 of
 * Alhajras/Raytracer-Data-structures. GPU_LBVH is appended so that the
 * existing values keep their numbers (BVH=0 ... NONE=4); a caller switching
 * on dataStructure (main.cpp:330, :791) gains one case.
 */
#ifndef RQ_SETTINGS_H
#define RQ_SETTINGS_H
#include <stdint.h>

#include "rq_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { RQ_BVH = 0, RQ_KDTREE = 1, RQ_UNIFORM_GRID = 2, RQ_LBVH = 3, RQ_NONE = 4,
               RQ_GPU_LBVH = 5 } rq_acc_type;

/* settings.h:5 */
typedef enum { RQ_IGEA, RQ_ARMADILLO, RQ_BUNNY, RQ_BUNNIES, RQ_TEST, RQ_GRASS, RQ_BUDDHA,
               RQ_CITY } rq_scene_model;

typedef struct {
  uint32_t width;          /* 640 */
  uint32_t height;         /* 480 */
  float fov;               /* 90 (the render loop hard-codes 30, main.cpp:545) */
  float background[3];     /* (1, 1, 1), unused by the render loop */
  float bias;              /* 0.0001 */
  uint32_t aa_samples;     /* 1 */
  rq_acc_type data_structure;
  rq_scene_model scene_model;
} rq_settings;

/* The reference defaults of settings.h:9-16 with data_structure = GPU_LBVH. */
static inline rq_settings rq_settings_default(void) {
  rq_settings s;
  s.width = 640;
  s.height = 480;
  s.fov = 90.0f;
  s.background[0] = s.background[1] = s.background[2] = 1.0f;
  s.bias = 0.0001f;
  s.aa_samples = 1;
  s.data_structure = RQ_GPU_LBVH;
  s.scene_model = RQ_BUNNY;
  return s;
}

/* Render parameters for these settings: image size and samples from the
 * settings, everything else the reference's constants (rq_render_defaults). */
static inline rq_render_params rq_settings_render_params(const rq_settings* s) {
  rq_render_params p;
  rq_render_defaults(&p);
  p.width = s->width;
  p.height = s->height;
  p.spp = s->aa_samples;
  return p;
}

#ifdef __cplusplus
}
#endif
#endif
