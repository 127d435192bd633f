// Internal types of the wavefront renderer (kernels <-> host orchestration).
#pragma once
#include "rq_internal.h"

namespace rq {

struct CamParams {
  uint32_t w, h, spp;
  float invW, invH, aspect, angle;  // as main.cpp:544-546 (angle = tan(pi*0.5*30/180) in float)
  uint64_t seed;
};

// Pinhole camera (serving requests): basis f (forward), r (right), u (up).
struct PinholeParams {
  float eye[3], f[3], r[3], u[3];
  float tan_half, aspect, inv_w, inv_h;
  uint32_t w, n;
  uint64_t seed;
};

struct ShadeParams {
  float light[3], emission[3], surface[3], sky[3];
  float bias;
  int32_t bounce;
  int32_t shadows;
  int32_t materials;  // castRay's material switch on SceneView::mat
};

// Triangles: vertices + indices. Spheres: centres in vx/vy/vz, radii in r.
// rgb: optional surface colour per primitive (else ShadeParams::surface).
struct SceneView {
  const float *vx, *vy, *vz;
  const uint32_t* idx;
  const float* r;
  const float* rgb;
  const uint8_t* mat;  // material per primitive, or null: all DIFFUSE_AND_GLOSSY
};

struct RaySoA {
  float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
};

hipError_t launch_raygen_camera(const LaunchCtx& lc, const CamParams& cp, const double* jit,
                                const RaySoA& R);
hipError_t launch_raygen_uniform(const LaunchCtx& lc, uint32_t n, uint64_t seed, const float lo[3],
                                 const float hi[3], const RaySoA& R);
hipError_t launch_raygen_pinhole(const LaunchCtx& lc, const PinholeParams& p, const RaySoA& R);
hipError_t launch_shade_primary(const LaunchCtx& lc, const SceneView& s, const ShadeParams& sp,
                                uint32_t n, const RaySoA& prim, const int32_t* h0, const float* t0,
                                const RaySoA& shad, const RaySoA& bnc, float4* lit0);
// Compacted copy of the second hits' shadow rays (all null: not kept):
// rays[k] with its sample idx[k], k < *count; occ: the second-hit flags,
// zeroed for the samples without a shadow ray.
struct CompactOut {
  RaySoA rays;
  uint32_t* idx;
  uint32_t* count;
  uint8_t* occ;
};
hipError_t launch_shade_bounce(const LaunchCtx& lc, const SceneView& s, const ShadeParams& sp,
                               uint32_t n, const RaySoA& bnc, const int32_t* h1, const float* t1,
                               const RaySoA& shad, float2* lit1, const CompactOut& co);
// Adds the number of active rays (tmin < tmax) of R[0, n) to *counter.
hipError_t launch_count_active(const LaunchCtx& lc, const RaySoA& R, uint64_t n,
                               uint32_t* counter);
// castRay's KDTREE case: per sample black when occ, else the sky; hit0 = -1.
hipError_t launch_shade_occlusion(const LaunchCtx& lc, const ShadeParams& sp, uint32_t npix,
                                  uint32_t spp, const uint8_t* occ, int32_t* h0, float* srgb,
                                  float* img);
hipError_t launch_shade_final(const LaunchCtx& lc, const SceneView& s, const ShadeParams& sp,
                              uint32_t npix, uint32_t spp, const RaySoA& prim, const RaySoA& bnc,
                              const int32_t* h0, const float* t0, const int32_t* h1,
                              const float* t1, const uint8_t* occ, const float4* lit0,
                              const float2* lit1, float* srgb, float* img);

}  // namespace rq
