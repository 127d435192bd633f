// Oracle-side wrapper around the reference's own sources (test
// infrastructure only). It is compiled by oracle/Makefile from the reference
// checkout where it lies (-I <reference>/project/raytracer) into
// oracle/_ref/libref.so, and exports C entry points that call the reference's
// functions unchanged. Nothing here restates an algorithm except the
// candidate loop, which repeats main.cpp:343-358 / :376-386 around the
// reference's own boxIntersect and Sphere::raySphereIntersect calls.
#include <memory>
#define main rq_reference_main
#include "main.cpp"
#undef main

extern "C" {

unsigned int ref_expand_bits(unsigned int v) { return expandBits(v); }

unsigned int ref_morton3d(float x, float y, float z) { return morton3D(x, y, z); }

// constructLBVHTree's code computation (accelerators.h:576-580) for centres.
void ref_morton_centers(const float* cx, const float* cy, const float* cz, unsigned n,
                        unsigned* codes) {
  for (unsigned i = 0; i < n; ++i) {
    Vec3f centroid = (Vec3f(cx[i], cy[i], cz[i]) + 30) / 1000;
    codes[i] = morton3D(centroid.x, centroid.y, centroid.z);
  }
}

int ref_ray_sphere(const float* c, float r, const float* o, const float* d, float* t0,
                   float* t1) {
  Sphere s(0, Vec3f(c[0], c[1], c[2]), r, Vec3f(0));
  return s.raySphereIntersect(Vec3f(o[0], o[1], o[2]), Vec3f(d[0], d[1], d[2]), *t0, *t1) ? 1 : 0;
}

int ref_box_hit(const float* lo, const float* hi, const float* o, const float* d) {
  auto node = std::make_shared<Node>();
  node->boxBoundries = BoxBoundries(Vec3f(lo[0], lo[1], lo[2]), Vec3f(hi[0], hi[1], hi[2]));
  return boundingBoxIntersection(Vec3f(o[0], o[1], o[2]), Vec3f(d[0], d[1], d[2]), node) ? 1 : 0;
}

// Sphere scene closest hit through the reference's BVH (accelerators.h:246
// build, :668 traversal) or, with use_bvh = 0, the NONE loop. Spheres are
// built exactly like createScene_new's objects (main.cpp:677-689). Returns
// the hit object id (-1 on miss) and t; also the BVH node count.
int ref_sphere_closest(const float* cx, const float* cy, const float* cz, const float* rad,
                       unsigned n, const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz, unsigned nrays,
                       int use_bvh, int* hit, float* tout) {
  sceneFixed.clear();
  std::vector<SceneObject> scene;
  for (unsigned i = 0; i < n; ++i) {
    SceneObject s;
    s.objId = (int)i;
    s.radius = rad[i];
    s.center = Vec3f(cx[i], cy[i], cz[i]);
    s.position = s.center;
    s.isSphere = true;
    s.boxBoundries = BoxBoundries(s.center - s.radius, s.center + s.radius);
    s.sphere = Sphere((int)i, s.center, s.radius, Vec3f(0.8, 0.7, 0), 0, 0.0);
    scene.push_back(s);
    sceneFixed.push_back(s);
  }
  int totalNodes = 0;
  std::shared_ptr<Node> root;
  if (use_bvh) root = constructBVHNew(scene, 0, (int)scene.size(), &totalNodes);
  for (unsigned r = 0; r < nrays; ++r) {
    Vec3f o(ox[r], oy[r], oz[r]), d(dx[r], dy[r], dz[r]);
    float tnear = INFINITY, t0, t1;
    int best = -1;
    if (use_bvh) {
      std::vector<int> cand;
      boxIntersect(o, d, root, cand);
      for (int box : cand) {
        t0 = INFINITY, t1 = INFINITY;
        if (sceneFixed[box].sphere.raySphereIntersect(o, d, t0, t1)) {
          if (t0 < 0) t0 = t1;
          if (t0 < tnear) tnear = t0, best = box;
        }
      }
    } else {
      for (unsigned i = 0; i < scene.size(); ++i) {
        t0 = INFINITY, t1 = INFINITY;
        if (sceneFixed[i].sphere.raySphereIntersect(o, d, t0, t1)) {
          if (t0 < 0) t0 = t1;
          if (t0 < tnear) tnear = t0, best = (int)i;
        }
      }
    }
    hit[r] = best;
    tout[r] = tnear;
  }
  return totalNodes;
}

// castRay itself (main.cpp:291-500) over a sphere scene built like
// createScene_new's objects (main.cpp:677-689, per-sphere surface colour),
// with the render loop's light (main.cpp:775), the NONE data structure and
// depth 1 (main.cpp:558): one colour per ray.
// castRay (NONE loop) over a sphere set, with mat[i] (or all
// DIFFUSE_AND_GLOSSY when null) as each sphere's materialType: the switch of
// main.cpp:417-446, including its recursion to depth 2.
void ref_sphere_scene_colors_mat(const float* cx, const float* cy, const float* cz,
                                 const float* rad, const float* rgb, const unsigned char* mat,
                                 unsigned n, const float* ox, const float* oy, const float* oz,
                                 const float* dx, const float* dy, const float* dz,
                                 unsigned nrays, float* out);

void ref_sphere_scene_colors(const float* cx, const float* cy, const float* cz, const float* rad,
                             const float* rgb, unsigned n, const float* ox, const float* oy,
                             const float* oz, const float* dx, const float* dy, const float* dz,
                             unsigned nrays, float* out) {
  ref_sphere_scene_colors_mat(cx, cy, cz, rad, rgb, nullptr, n, ox, oy, oz, dx, dy, dz, nrays,
                              out);
}

void ref_sphere_scene_colors_mat(const float* cx, const float* cy, const float* cz,
                                 const float* rad, const float* rgb, const unsigned char* mat,
                                 unsigned n, const float* ox, const float* oy, const float* oz,
                                 const float* dx, const float* dy, const float* dz,
                                 unsigned nrays, float* out) {
  sceneFixed.clear();
  std::vector<SceneObject> scene;
  for (unsigned i = 0; i < n; ++i) {
    SceneObject s;
    s.objId = (int)i;
    s.radius = rad[i];
    s.center = Vec3f(cx[i], cy[i], cz[i]);
    s.position = s.center;
    s.shininess = 64;
    s.isSphere = true;
    s.boxBoundries = BoxBoundries(s.center - s.radius, s.center + s.radius);
    s.sphere = Sphere((int)i, s.center, s.radius, Vec3f(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]),
                      0, 0.0);
    if (mat) s.sphere.materialType = (MaterialType)mat[i];
    scene.push_back(s);
    sceneFixed.push_back(s);
  }
  std::vector<Sphere> spheres;
  std::vector<Sphere> lights{Sphere(0, Vec3f(0, 3, 30), 10, Vec3f(1, 1, 1), 0, 0.0, Vec3f(1))};
  Settings settings;
  settings.dataStructure = NONE;
  std::shared_ptr<Node> root;
  const int depth = 1;
  for (unsigned r = 0; r < nrays; ++r) {
    Vec3f c = castRay(Vec3f(ox[r], oy[r], oz[r]), Vec3f(dx[r], dy[r], dz[r]), spheres, lights,
                      scene, root, depth, settings);
    out[3 * r] = c.x, out[3 * r + 1] = c.y, out[3 * r + 2] = c.z;
  }
}

// The reference's KDTREE path: its own SAH kd-tree over the spheres
// (constructKDTreeNew with main.cpp:815-819's arguments) and kdtreeIntersect
// per ray (accelerators.h:996-1086). The tree lives in the reference's
// globals, reset here to their initial state before the build.
void ref_kdtree_any(const float* cx, const float* cy, const float* cz, const float* rad,
                    unsigned n, const float* ox, const float* oy, const float* oz,
                    const float* dx, const float* dy, const float* dz, unsigned nrays,
                    unsigned char* out) {
  std::vector<SceneObject> scene;
  for (unsigned i = 0; i < n; ++i) {
    SceneObject s;
    s.objId = (int)i;
    s.radius = rad[i];
    s.center = Vec3f(cx[i], cy[i], cz[i]);
    s.position = s.center;
    s.shininess = 64;
    s.isSphere = true;
    s.boxBoundries = BoxBoundries(s.center - s.radius, s.center + s.radius);
    s.sphere = Sphere((int)i, s.center, s.radius, Vec3f(0.8, 0.7, 0), 0, 0.0);
    scene.push_back(s);
  }
  nAllocedNodes = 0, nextFreeNode = 0, nodes = nullptr, totalKdNodes = 0;
  bounds = BoxBoundries();
  primBounds.clear();
  kdtreePrimitiveIndices.clear();
  constructKDTreeNew(scene, 80, 1, 0.5f, 1, -1);
  for (unsigned r = 0; r < nrays; ++r)
    out[r] = kdtreeIntersect(Vec3f(ox[r], oy[r], oz[r]), Vec3f(dx[r], dy[r], dz[r])) ? 1 : 0;
}

// render()'s primary ray directions (main.cpp:542-556) through the
// reference's own Vec3f::normalize, for given jitter pairs.
void ref_camera_dirs(unsigned w, unsigned h, unsigned spp, const double* jitter, float* dx,
                     float* dy, float* dz) {
  float invWidth = 1 / float(w), invHeight = 1 / float(h);
  float fov = 30, aspectratio = w / float(h);
  float angle = tan(M_PI * 0.5 * fov / 180.);
  size_t i = 0;
  for (unsigned y = 0; y < h; ++y)
    for (unsigned x = 0; x < w; ++x)
      for (unsigned k = 0; k < spp; ++k, ++i) {
        float xx = (2 * ((x + jitter[2 * i]) * invWidth) - 1) * angle * aspectratio;
        float yy = (1 - 2 * ((y + jitter[2 * i + 1]) * invHeight)) * angle;
        Vec3f raydir(xx, yy, -1);
        raydir.normalize();
        dx[i] = raydir.x, dy[i] = raydir.y, dz[i] = raydir.z;
      }
}

// The reference's mt19937 + uniform_real_distribution jitter stream
// (main.cpp:503-508): n successive random_double() values.
void ref_random_double(unsigned n, double* out) {
  for (unsigned i = 0; i < n; ++i) out[i] = random_double();
}

}  // extern "C"
