/*
 * rq — host-side scene generation and OBJ I/O (C ABI).
 *
 * Replaces createScene_new (main.cpp:599-721), which reads only `v` lines and
 * turns each vertex into a sphere. Here a mesh is SoA vertices plus a
 * 0-based triangle index list, ready for rq_scene_upload. All generators are
 * deterministic functions of their arguments (integer hashing, no libc RNG).
 * Meshes returned by the library are freed with rq_mesh_free.
 */
#ifndef RQ_SCENE_H
#define RQ_SCENE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float *vx, *vy, *vz;
  uint32_t nv;
  uint32_t* idx; /* 3 per triangle */
  uint32_t ntri;
} rq_mesh;

/* Icosahedron subdivided `level` times (midpoints projected to the unit
 * sphere; V = 10*4^level + 2, F = 20*4^level), each vertex p displaced to
 * center + scale * (1 + amp * vnoise(freq * p, seed)) * p, where vnoise is
 * trilinear value noise in [-1, 1) on the integer lattice (smoothstep weights,
 * lattice values from a 32-bit integer hash of (ix, iy, iz, seed)). */
int rq_gen_icosphere(uint32_t level, uint32_t seed, float amp, float freq, const float center[3],
                     float scale, rq_mesh* out);

/* `count` displaced icospheres of `level`, seeds seed0 + i, centres drawn by
 * dart throwing (minimum distance rmin) in [0, box]^3 from a splitmix64 stream
 * keyed by placement_seed, merged into one mesh in sphere order. */
int rq_gen_icosphere_field(uint32_t count, uint32_t level, uint32_t seed0, uint64_t placement_seed,
                           float box, float rmin, float amp, float freq, rq_mesh* out);

/* `ntri` independent triangles: centroid uniform in the unit cube, random
 * orientation, edge length ~ N(mean, sd) clipped to >= min_edge. */
int rq_gen_soup(uint32_t ntri, uint64_t seed, float mean_edge, float sd_edge, float min_edge,
                rq_mesh* out);

/* Wavefront OBJ: `v x y z` and `f a b c ...` (1-based or negative indices,
 * `a/b/c` forms, polygons fanned into triangles); other lines ignored. */
int rq_obj_read(const char* path, rq_mesh* out);
/* The same reader with an explicit thread count (0 = up to 16, 1 = one pass on
 * the calling thread): the file is memory-mapped and parsed in line-aligned
 * chunks, relative indices resolved against the vertices before each record;
 * the result is identical for every thread count. */
int rq_obj_read_ex(const char* path, int threads, rq_mesh* out);
int rq_obj_write(const char* path, const rq_mesh* m);

/* Spheres (host SoA): centres, radii and a surface colour (3 floats) each. */
typedef struct {
  float *cx, *cy, *cz, *r, *rgb;
  uint32_t n;
} rq_spheres;

/* The reference's scene (createScene_new, main.cpp:599-721): the `v` records
 * at the start of an OBJ file, read up to the first record of another kind
 * (or `max_vertices`, 0 = no limit), each one a sphere of radius 0.05 at
 * v * 100 + clone * 20, then y - 10, z - 60 (float arithmetic in that order),
 * colour (0.8, 0.7, 0), repeated for `clones` copies of the file (the
 * reference's NUMBER_OF_CLONES, 1); then the ground sphere of radius 100 at
 * (0.93591022, -105.47120094, -43.2363205), colour (0, 0, 0). */
int rq_gen_reference_scene(const char* obj_path, uint32_t clones, uint32_t max_vertices,
                           rq_spheres* out);
void rq_spheres_free(rq_spheres* s);

/* Axis-aligned bounds of the vertices (lo xyz, hi xyz). */
void rq_mesh_bounds(const rq_mesh* m, float lo[3], float hi[3]);
void rq_mesh_free(rq_mesh* m);

#ifdef __cplusplus
}
#endif
#endif
