/*
 * CPU ORACLE — test infrastructure only.
 *
 * This library is the oracle the GPU ray-query engine is tested against. It is a plain
 * CPU restatement of the reference algorithms on the acceleration-structure
 * path of Alhajras/Raytracer-Data-structures (project/raytracer/...). Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library never links, calls or falls back to it.
 *
 * Every function cites the reference line(s) it restates. Where the reference
 * has no live implementation (triangles, Karras emission, shadow rays) the
 * function restates the intended algorithm named at the cited lines and is
 * pinned by closed forms, independent algorithms and brute force in tests/.
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   or_expand_bits / or_morton3d   exhaustive + compiled reference (oracle/_ref)
 *   or_morton_codes                closed-form interleave + compiled reference
 *   or_sort_pairs                  std::stable_sort invariants + numpy lexsort
 *   or_karras                      independent top-down builder + brute split
 *   or_refit                       recursive union + containment invariants
 *   or_mt                          closed-form known answers (axis triangles)
 *   or_closest_* / or_any_*        brute force with lowest-index tie rule
 *   or_ray_sphere                  compiled reference Sphere::raySphereIntersect
 *   or_*_brute_spheres             compiled reference NONE loop (castRay, fixture + live)
 *   or_bvh_build_spheres           box-filtered brute force; reference castRay colours
 *   normalize (camera, shading)    the reference's Vec3f::normalize and render() camera
 *   or_mt19937_jitter              C++ standard known answer + numpy MT19937
 *   or_philox4x32_10               Random123 published known-answer vectors
 *   or_shade_*                     closed-form single-light cases
 *   or_treelet_dp                  brute force over every binary tree of 2..7 leaves
 *   or_bvh_treelets                invariants (leaf set, box unions, cost and height
 *                                  never grow), queries vs brute force
 */
#ifndef RQ_ORACLE_H
#define RQ_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- Morton (accelerators.h:374-394, normalisation :577) ---- */
uint32_t or_expand_bits(uint32_t v);
uint32_t or_morton3d(float x, float y, float z);
/* mode 0 = REF_AFFINE ((c+30)/1000), mode 1 = SCENE_AABB (centroid bounds). */
int or_morton_codes(const float* vx, const float* vy, const float* vz, uint32_t nv,
                    const uint32_t* idx, uint32_t ntri, int mode, uint32_t* codes);
void or_centroid(const float* vx, const float* vy, const float* vz,
                 const uint32_t* idx, uint32_t tri, float c[3]);

/* ---- sort + topology (accelerators.h:348-369 intent; :397-449 intent) ---- */
void or_sort_pairs(const uint32_t* codes, uint32_t n, uint32_t* sorted_codes, uint32_t* sorted_idx);
/* Internal node i in [0,n-2]; child >= 0 is an internal node, child < 0 is leaf ~child. */
void or_karras(const uint32_t* sorted_codes, uint32_t n, int32_t* left, int32_t* right);
void or_topdown(const uint32_t* sorted_codes, uint32_t n, int32_t* left, int32_t* right);
/* boxes: (n-1) internal boxes then n leaf boxes, 6 floats each (lo xyz, hi xyz). */
void or_refit(const float* vx, const float* vy, const float* vz, const uint32_t* idx,
              const uint32_t* sorted_idx, uint32_t n, const int32_t* left, const int32_t* right,
              float* boxes);
uint32_t or_tree_height(const int32_t* left, const int32_t* right, uint32_t n);

/* ---- triangle test (main.cpp:137-162, fixed v_0 -> v0, plus t in (tmin,tmax)) ---- */
int or_mt(const float v0[3], const float v1[3], const float v2[3],
          const float o[3], const float d[3], float tmin, float tmax,
          float* t, float* u, float* v);

/* ---- sphere test (accelerators.h:79-92) ---- */
int or_ray_sphere(const float c[3], float r, const float o[3], const float d[3], float* t0, float* t1);

/* ---- queries ---- */
typedef struct {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} or_rays;
typedef struct {
  int32_t* prim;
  float *t, *u, *v;
} or_hits;

/* Brute force (main.cpp:373-387 NONE loop): lowest index wins ties. */
void or_closest_brute(const float* vx, const float* vy, const float* vz, const uint32_t* idx,
                      uint32_t ntri, const or_rays* rays, uint32_t nrays, or_hits* hits, int nthreads);
void or_any_brute(const float* vx, const float* vy, const float* vz, const uint32_t* idx,
                  uint32_t ntri, const or_rays* rays, uint32_t nrays, uint8_t* occ, int nthreads);

/* NONE loop over spheres (main.cpp:376-386) with the sphere leaf of castRay:
 * t0, or t1 when t0 < 0 (main.cpp:349-353), accepted in [tmin, tmax).
 * box_filter: only spheres whose box the ray hits (the reference BVH path's
 * candidates, main.cpp:343-358), the semantics of any box hierarchy. */
void or_closest_brute_spheres(const float* cx, const float* cy, const float* cz, const float* r,
                              uint32_t n, const or_rays* rays, uint32_t nrays, or_hits* hits,
                              int box_filter, int nthreads);
void or_any_brute_spheres(const float* cx, const float* cy, const float* cz, const float* r,
                          uint32_t n, const or_rays* rays, uint32_t nrays, uint8_t* occ,
                          int nthreads);

/* Canonical LBVH traversal over an oracle-built tree. visits/tests may be NULL. */
typedef struct or_bvh or_bvh;
or_bvh* or_bvh_build(const float* vx, const float* vy, const float* vz, uint32_t nv,
                     const uint32_t* idx, uint32_t ntri, int morton_mode);
/* Sphere scene (accelerators.h:54-92 objects): centroid = centre, leaf box
 * centre -/+ r (main.cpp:687-689), the same Morton / sort / Karras / refit. */
or_bvh* or_bvh_build_spheres(const float* cx, const float* cy, const float* cz, const float* r,
                             uint32_t n, int morton_mode);
void or_bvh_free(or_bvh* b);
void or_bvh_arrays(const or_bvh* b, uint32_t* codes, uint32_t* sorted_idx, int32_t* left,
                   int32_t* right, float* boxes);
uint32_t or_bvh_height(const or_bvh* b);
/* Treelet restructuring of the tree in place (one round of 7-leaf treelets,
 * Karras & Aila 2013); returns the number of treelets replaced. */
uint32_t or_bvh_treelets(or_bvh* b);
/* The optimal tree over m <= 7 boxes (cost = sum of internal half areas):
 * its cost, and per subset s the split (the part with the lowest leaf) and
 * the height (leaves at height 0). */
float or_treelet_dp(const float* boxes, int m, uint8_t* split, uint8_t* height);
void or_closest_bvh(const or_bvh* b, const or_rays* rays, uint32_t nrays, or_hits* hits,
                    uint32_t* visits, uint32_t* tests, int nthreads);
/* Deepest stack of each ray's canonical closest-hit walk. */
void or_closest_bvh_stack(const or_bvh* b, const or_rays* rays, uint32_t nrays,
                          uint32_t* max_stack, int nthreads);
void or_any_bvh(const or_bvh* b, const or_rays* rays, uint32_t nrays, uint8_t* occ,
                uint32_t* visits, uint32_t* tests, int nthreads);

/* ---- random numbers and rays ---- */
void or_mt19937_jitter(uint32_t n, double* out);            /* main.cpp:503-508 */
void or_mt19937_raw(uint32_t n, uint32_t* out);
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* main.cpp:541-558: reference pinhole, jitter pairs given (2 doubles per sample). */
void or_camera_rays_ref(uint32_t w, uint32_t h, uint32_t spp, const double* jitter,
                        float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                        float* tmin, float* tmax);
void or_camera_rays_philox(uint32_t w, uint32_t h, uint32_t spp, uint64_t seed,
                           float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                           float* tmin, float* tmax);
void or_uniform_rays(uint32_t n, uint64_t seed, const float lo[3], const float hi[3],
                     float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                     float* tmin, float* tmax);
/* Rays ids[k] (k < n; ids NULL: rays 0..n-1) of or_uniform_rays, written at k. */
void or_uniform_rays_at(const uint32_t* ids, uint32_t n, uint64_t seed, const float lo[3],
                        const float hi[3], float* ox, float* oy, float* oz, float* dx, float* dy,
                        float* dz, float* tmin, float* tmax);
/* Rays ids[k] (k < count; ids NULL: rays 0..count-1) of the serving pinhole
   camera for a request of n rays, written at k. */
void or_pinhole_rays_at(const uint32_t* ids, uint32_t count, uint32_t n, uint64_t seed,
                        const float eye[3], const float look_at[3], const float up[3],
                        float fov_deg, uint32_t width, float* ox, float* oy, float* oz, float* dx,
                        float* dy, float* dz, float* tmin, float* tmax);

/* ---- shading (main.cpp:394-497, fresnel :85-103, reflect :218-221) ---- */
typedef struct {
  float light[3];
  float emission[3];
  float surface[3];
  float sky[3];
  float bias;
  int bounce;  /* 0: primary + shadow (Lambert config); 1: + mirror bounce */
  int shadows; /* 1: shadow ray per hit; 0: trace_more stub (main.cpp:235-245), never shadowed */
  const float* prim_rgb; /* surface colour per primitive (3 floats), or NULL for `surface` */
  const uint8_t* prim_mat; /* material per primitive (accelerators.h:24: 0 DIFFUSE_AND_GLOSSY,
                              1 REFLECTION_AND_REFRACTION, 2 REFLECTION), or NULL: all 0 */
  int mode; /* 0: castRay shading; 1: castRay's KDTREE case (main.cpp:362-370): per sample
               black when the ray hits anything (any hit), else the sky */
} or_shade_params;
/* Full frame on the CPU: rays given, returns per-sample colour (3 floats) and the
   primary/bounce hits plus the shadow booleans. Any output pointer may be NULL. */
void or_render_samples(const or_bvh* b, const float* vx, const float* vy, const float* vz,
                       const uint32_t* idx, const or_rays* primary, uint32_t nrays,
                       const or_shade_params* sp, float* rgb, int32_t* hit0, int32_t* hit1,
                       uint8_t* shadow0, uint8_t* shadow1, uint64_t* counts, int nthreads);
/* main.cpp:516-528: (uchar)(min(1, c/aa)*255) after summing spp samples per pixel. */
void or_resolve_ppm(const float* rgb, uint32_t npix, uint32_t spp, uint8_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
